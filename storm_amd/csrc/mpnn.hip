// MI355X (gfx950) kernels and C ABI of the warm-start network
// (include/storm_mpnn.h): an edge message-passing network whose every
// product is a BF16 MFMA GEMM (v_mfma_f32_16x16x16_bf16) with FP32
// accumulation, trained with AdamW.  Activations and gradients are stored in
// FP32 and rounded to BF16 (nearest even) as they enter a product.  Sums
// over edges and over rows go through fixed orders (CSR gathers, per-block
// partials reduced in block order): no atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/storm_mpnn.h"

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int HID = 64;      // hidden width
constexpr int ROWS_PER_PART = 1024;  // rows per partial of a weight gradient (general K)
constexpr int DW_ROWS = 128;         // rows per partial of k_gemm_dw_full (one wave each)
constexpr int RED_CHUNK = 32;        // partials summed per thread in the first reduction level

__device__ __forceinline__ short to_bf16(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (short)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// 16x16x16 product: lane l holds A[l%16][4*(l/16)+j], B[4*(l/16)+j][l%16]
// and returns D[4*(l/16)+j][l%16], j = 0..3.
__device__ __forceinline__ v4f mfma16(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Y[M][NOUT] = act(X[M][K] W[K][NOUT] + b): one wave per 16 rows.
template <int NOUT>
__global__ void __launch_bounds__(256)
k_gemm(const float* X, const float* W, const float* b, int M, int K, int relu, float* Y) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r0 = (blockIdx.x * 4 + wave) * 16;
  if (r0 >= M) return;  // whole wave
  constexpr int NT = NOUT / 16;
  v4f acc[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) acc[nt] = v4f{0.f, 0.f, 0.f, 0.f};
  const int ar = r0 + (l & 15), kq = 4 * (l >> 4), cn = l & 15;
  for (int k0 = 0; k0 < K; k0 += 16) {
    v4s a;
#pragma unroll
    for (int j = 0; j < 4; ++j) a[j] = (ar < M) ? to_bf16(X[(size_t)ar * K + k0 + kq + j]) : (short)0;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      v4s bb;
#pragma unroll
      for (int j = 0; j < 4; ++j) bb[j] = to_bf16(W[(size_t)(k0 + kq + j) * NOUT + nt * 16 + cn]);
      acc[nt] = mfma16(a, bb, acc[nt]);
    }
  }
#pragma unroll
  for (int nt = 0; nt < NT; ++nt)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = r0 + kq + j, col = nt * 16 + cn;
      if (row < M) {
        float v = acc[nt][j] + b[col];
        if (relu) v = fmaxf(v, 0.f);
        Y[(size_t)row * NOUT + col] = v;
      }
    }
}

// dX[M][K] = dZ[M][NOUT] W^T, dZ = dY masked where act <= 0 (act may be null).
template <int NOUT>
__global__ void __launch_bounds__(256)
k_gemm_dx(const float* dY, const float* act, const float* W, int M, int K, float* dX) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r0 = (blockIdx.x * 4 + wave) * 16;
  if (r0 >= M) return;
  constexpr int NK = NOUT / 16;
  const int ar = r0 + (l & 15), kq = 4 * (l >> 4), cn = l & 15;
  v4s a[NK];
#pragma unroll
  for (int q = 0; q < NK; ++q)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float v = 0.f;
      if (ar < M) {
        const size_t ix = (size_t)ar * NOUT + q * 16 + kq + j;
        v = dY[ix];
        if (act && !(act[ix] > 0.f)) v = 0.f;
      }
      a[q][j] = to_bf16(v);
    }
  for (int kt = 0; kt < K / 16; ++kt) {
    v4f acc = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < NK; ++q) {
      v4s bb;  // B[n][k] = W[k][n]
#pragma unroll
      for (int j = 0; j < 4; ++j) bb[j] = to_bf16(W[(size_t)(kt * 16 + cn) * NOUT + q * 16 + kq + j]);
      acc = mfma16(a[q], bb, acc);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = r0 + kq + j;
      if (row < M) dX[(size_t)row * K + kt * 16 + cn] = acc[j];
    }
  }
}

// Partial weight gradient of rows [blockIdx.x*ROWS_PER_PART, +ROWS_PER_PART):
// part[blk][K][NOUT] = X^T dZ.  The block's 4 waves share the (K/16)x(NOUT/16)
// tiles; rows beyond M count as zero.
template <int NOUT>
__global__ void __launch_bounds__(256)
k_gemm_dw(const float* X, const float* dY, const float* act, int M, int K, float* part) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  constexpr int NT = NOUT / 16;
  const int ntiles = (K / 16) * NT;
  const int m_begin = blockIdx.x * ROWS_PER_PART;
  const int m_end = min(M, m_begin + ROWS_PER_PART);
  const int kq = 4 * (l >> 4), cn = l & 15;
  float* out = part + (size_t)blockIdx.x * K * NOUT;
  for (int tile = wave; tile < ntiles; tile += 4) {
    const int kt = tile / NT, nt = tile % NT;
    v4f acc = v4f{0.f, 0.f, 0.f, 0.f};
    for (int m = m_begin; m < m_end; m += 16) {
      v4s a, bb;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = m + kq + j;
        float x = 0.f, d = 0.f;
        if (row < m_end) {
          x = X[(size_t)row * K + kt * 16 + cn];
          const size_t iy = (size_t)row * NOUT + nt * 16 + cn;
          d = dY[iy];
          if (act && !(act[iy] > 0.f)) d = 0.f;
        }
        a[j] = to_bf16(x);   // A[k][m] = X[m][k]
        bb[j] = to_bf16(d);  // B[m][n] = dZ[m][n]
      }
      acc = mfma16(a, bb, acc);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) out[(size_t)(kt * 16 + kq + j) * NOUT + nt * 16 + cn] = acc[j];
  }
}

// Partial weight and bias gradients of rows [blockIdx.x*DW_ROWS, +DW_ROWS)
// for K = 16*KT, one wave per block: part[blk][K][NOUT] = X^T dZ and
// partb[blk][n] = the sum of dZ[m][n] (FP32, unrounded).  The wave holds every
// (k, n) tile of the product, so X and dY are read once.  Rows beyond M count
// as zero.
template <int NOUT, int KT>
__global__ void __launch_bounds__(64)
k_gemm_dw_full(const float* X, const float* dY, const float* act, int M, float* part,
               float* partb) {
  constexpr int K = 16 * KT, NT = NOUT / 16;
  __shared__ float s_b[4][NOUT];
  const int l = threadIdx.x;
  const int kq = 4 * (l >> 4), cn = l & 15;
  v4f acc[KT][NT];
#pragma unroll
  for (int kt = 0; kt < KT; ++kt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[kt][nt] = v4f{0.f, 0.f, 0.f, 0.f};
  float cs[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) cs[nt] = 0.f;
  const int w_begin = blockIdx.x * DW_ROWS;
  const int w_end = min(M, w_begin + DW_ROWS);
  for (int m = w_begin; m < w_end; m += 16) {
    v4s b[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = m + kq + j;
        float d = 0.f;
        if (row < w_end) {
          const size_t iy = (size_t)row * NOUT + nt * 16 + cn;
          d = dY[iy];
          if (act && !(act[iy] > 0.f)) d = 0.f;
        }
        cs[nt] += d;
        b[nt][j] = to_bf16(d);  // B[m][n] = dZ[m][n]
      }
#pragma unroll
    for (int kt = 0; kt < KT; ++kt) {
      v4s a;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = m + kq + j;
        a[j] = to_bf16(row < w_end ? X[(size_t)row * K + kt * 16 + cn] : 0.f);  // A[k][m] = X[m][k]
      }
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[kt][nt] = mfma16(a, b[nt], acc[kt][nt]);
    }
  }
  float* out = part + (size_t)blockIdx.x * K * NOUT;
#pragma unroll
  for (int kt = 0; kt < KT; ++kt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int j = 0; j < 4; ++j) out[(size_t)(kt * 16 + kq + j) * NOUT + nt * 16 + cn] = acc[kt][nt][j];
  // column sums: the four row groups of each column, in group order
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) s_b[kq / 4][nt * 16 + cn] = cs[nt];
  __syncthreads();
  for (int n = l; n < NOUT; n += 64)
    partb[(size_t)blockIdx.x * NOUT + n] = ((s_b[0][n] + s_b[1][n]) + s_b[2][n]) + s_b[3][n];
}

// First reduction level: part2[c][i] = sum over blocks c*RED_CHUNK ..
// min(nblk, (c+1)*RED_CHUNK) - 1, in block order, of part[blk][i].
__global__ void k_reduce_chunk(const float* part, int nblk, int size, float* part2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= size) return;
  const int c = blockIdx.y, b0 = c * RED_CHUNK, b1 = min(nblk, b0 + RED_CHUNK);
  float s = 0.f;
  for (int b = b0; b < b1; ++b) s += part[(size_t)b * size + i];
  part2[(size_t)c * size + i] = s;
}

// Partial bias gradient: partb[blk][n] = sum of dZ[m][n] over the block's rows.
__global__ void k_colsum(const float* dY, const float* act, int M, int NOUT, float* partb) {
  const int n = threadIdx.x;
  if (n >= NOUT) return;
  const int m_begin = blockIdx.x * ROWS_PER_PART;
  const int m_end = min(M, m_begin + ROWS_PER_PART);
  float s = 0.f;
  for (int m = m_begin; m < m_end; ++m) {
    const size_t iy = (size_t)m * NOUT + n;
    float d = dY[iy];
    if (act && !(act[iy] > 0.f)) d = 0.f;
    s += d;
  }
  partb[(size_t)blockIdx.x * NOUT + n] = s;
}

// out[i] = sum over blocks, in block order, of part[blk][i].
__global__ void k_reduce(const float* part, int nblk, int size, float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= size) return;
  float s = 0.f;
  for (int b = 0; b < nblk; ++b) s += part[(size_t)b * size + i];
  out[i] = s;
}

// Xe[e] = [h_src, h_dst, ef[e]]
__global__ void k_edge_in(const float* H, const float* ef, const int* src, const int* dst, int E,
                          int fe, float* Xe) {
  const int K = 2 * HID + fe;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)E * K) return;
  const int e = (int)(t / K), k = (int)(t % K);
  float v;
  if (k < HID) v = H[(size_t)src[e] * HID + k];
  else if (k < 2 * HID) v = H[(size_t)dst[e] * HID + k - HID];
  else v = ef[(size_t)e * fe + k - 2 * HID];
  Xe[t] = v;
}

// Xn[i] = [h_i, sum of m_e over the edges into i (ascending edge ids)]
__global__ void k_node_in(const float* H, const float* Mg, const int* in_ptr, const int* in_idx,
                          int N, float* Xn) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)N * HID) return;
  const int i = (int)(t / HID), k = (int)(t % HID);
  float s = 0.f;
  for (int q = in_ptr[i]; q < in_ptr[i + 1]; ++q) s += Mg[(size_t)in_idx[q] * HID + k];
  Xn[(size_t)i * 2 * HID + k] = H[t];
  Xn[(size_t)i * 2 * HID + HID + k] = s;
}

__global__ void k_add(const float* a, const float* b, size_t n, float* out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = a[t] + b[t];
}

// dM[e] = dA[dst[e]] (the aggregate's gradient); dA is the second half of dXn
__global__ void k_edge_grad(const float* dXn, const int* dst, int E, float* dM) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)E * HID) return;
  const int e = (int)(t / HID), k = (int)(t % HID);
  dM[t] = dXn[(size_t)dst[e] * 2 * HID + HID + k];
}

// dH[i] = dHin[i] + dXn[i][0:H] + sum over edges out of i of dXe[e][0:H]
//       + sum over edges into i of dXe[e][H:2H]   (ascending edge ids)
__global__ void k_node_grad(const float* dHin, const float* dXn, const float* dXe, int Ke,
                            const int* out_ptr, const int* out_idx, const int* in_ptr,
                            const int* in_idx, int N, float* dH) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)N * HID) return;
  const int i = (int)(t / HID), k = (int)(t % HID);
  float s = dHin[t] + dXn[(size_t)i * 2 * HID + k];
  for (int q = out_ptr[i]; q < out_ptr[i + 1]; ++q) s += dXe[(size_t)out_idx[q] * Ke + k];
  for (int q = in_ptr[i]; q < in_ptr[i + 1]; ++q) s += dXe[(size_t)in_idx[q] * Ke + HID + k];
  dH[t] = s;
}

// loss terms and dY = 2 w (y - t) / (N F); lossp[blk] = partial sums
__global__ void __launch_bounds__(256)
k_loss(const float* Y, const float* T, const float* w, int N, int F, float* dY, double* lossp) {
  __shared__ double s[256];
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (t < (size_t)N * F) {
    const int i = (int)(t / F);
    const float wi = w ? w[i] : 1.f;
    const float d = Y[t] - T[t];
    acc = (double)wi * (double)d * (double)d;
    dY[t] = 2.f * wi * d / (float)((double)N * F);
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) lossp[blockIdx.x] = s[0];
}

__global__ void k_adamw(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                        float b2, float eps, float wd, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / c1, vh = vi / c2;
  p[i] = p[i] - lr * (mh / (sqrtf(vh) + eps) + wd * p[i]);
}

struct Err {
  std::string msg;
};

#define HIPCHK(x)                                                                 \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) throw Err{std::string(#x) + ": " + hipGetErrorString(e_)}; \
  } while (0)

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  void need(size_t k) {
    if (k <= n) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    HIPCHK(hipMalloc(&p, std::max<size_t>(k, 1) * sizeof(T)));
    n = k;
  }
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
};

inline unsigned blocks(size_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

}  // namespace

struct storm_mpnn {
  storm_mpnn_cfg_t cfg{};
  std::string err;
  hipStream_t stream = nullptr;
  int64_t np = 0;
  int64_t o_enc_w = 0, o_enc_b = 0, o_dec_w = 0, o_dec_b = 0;
  std::vector<int64_t> o_ew, o_eb, o_nw, o_nb;
  DBuf<float> p, g, m, v;
  int step = 0;
  int N = 0, E = 0;
  DBuf<int> src, dst, in_ptr, in_idx, out_ptr, out_idx;
  DBuf<float> xin, ef;
  std::vector<DBuf<float>> H, Xe, Mg, Xn, U;
  DBuf<float> Y, dY, dH, dH2, dXn, dM, dXe, dZ0, part, tgt, wgt;
  DBuf<double> lossp;
  bool have_fwd = false;

  int Ke() const { return 2 * HID + cfg.f_edge; }
  ~storm_mpnn() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
  }
};

namespace {

// Weight gradient of Y = act(X W + b) (rows M, inputs K, NOUT outputs): the
// partials of each row block, reduced in block order into gw and gb.
// Sum of the nblk rows of part[nblk][size] into out[size], in a fixed tree:
// chunks of RED_CHUNK rows in order, then the chunks in order.
void reduce_rows(storm_mpnn* S, const float* part, int nblk, int size, float* scratch, float* out) {
  const int nch = (nblk + RED_CHUNK - 1) / RED_CHUNK;
  if (nch <= 1) {
    hipLaunchKernelGGL(k_reduce, dim3(blocks(size)), dim3(256), 0, S->stream, part, nblk, size, out);
    return;
  }
  hipLaunchKernelGGL(k_reduce_chunk, dim3(blocks(size), nch), dim3(256), 0, S->stream, part, nblk,
                     size, scratch);
  hipLaunchKernelGGL(k_reduce, dim3(blocks(size)), dim3(256), 0, S->stream, scratch, nch, size, out);
}

template <int NOUT, int KT>
void weight_grad_full(storm_mpnn* S, const float* X, const float* dY, const float* act, int M,
                      float* gw, float* gb) {
  constexpr int K = 16 * KT;
  const int nblk = (M + DW_ROWS - 1) / DW_ROWS;
  const int nch = (nblk + RED_CHUNK - 1) / RED_CHUNK;
  S->part.need((size_t)(nblk + nch) * (K * NOUT + NOUT));
  float* pw = S->part.p;
  float* pb = pw + (size_t)nblk * K * NOUT;
  float* sw = pb + (size_t)nblk * NOUT;
  float* sb = sw + (size_t)nch * K * NOUT;
  hipLaunchKernelGGL((k_gemm_dw_full<NOUT, KT>), dim3(nblk), dim3(64), 0, S->stream, X, dY, act, M,
                     pw, pb);
  reduce_rows(S, pw, nblk, K * NOUT, sw, gw);
  reduce_rows(S, pb, nblk, NOUT, sb, gb);
}

template <int NOUT>
void weight_grad(storm_mpnn* S, const float* X, const float* dY, const float* act, int M, int K,
                 float* gw, float* gb) {
  // the widths of this network: every (k, n) tile held by one wave
  if (K == 16) return weight_grad_full<NOUT, 1>(S, X, dY, act, M, gw, gb);
  if (K == 64) return weight_grad_full<NOUT, 4>(S, X, dY, act, M, gw, gb);
  if (K == 128) return weight_grad_full<NOUT, 8>(S, X, dY, act, M, gw, gb);
  if (K == 144) return weight_grad_full<NOUT, 9>(S, X, dY, act, M, gw, gb);
  const int nblk = (M + ROWS_PER_PART - 1) / ROWS_PER_PART;
  S->part.need((size_t)nblk * (K * NOUT + NOUT));
  float* pw = S->part.p;
  float* pb = S->part.p + (size_t)nblk * K * NOUT;
  hipLaunchKernelGGL(k_gemm_dw<NOUT>, dim3(nblk), dim3(256), 0, S->stream, X, dY, act, M, K, pw);
  hipLaunchKernelGGL(k_colsum, dim3(nblk), dim3(64), 0, S->stream, dY, act, M, NOUT, pb);
  hipLaunchKernelGGL(k_reduce, dim3(blocks((size_t)K * NOUT)), dim3(256), 0, S->stream, pw, nblk,
                     K * NOUT, gw);
  hipLaunchKernelGGL(k_reduce, dim3(blocks(NOUT)), dim3(256), 0, S->stream, pb, nblk, NOUT, gb);
}

template <int NOUT>
void gemm(storm_mpnn* S, const float* X, const float* W, const float* b, int M, int K, int relu,
          float* Y) {
  hipLaunchKernelGGL(k_gemm<NOUT>, dim3(blocks(M, 64)), dim3(256), 0, S->stream, X, W, b, M, K,
                     relu, Y);
}

template <int NOUT>
void gemm_dx(storm_mpnn* S, const float* dY, const float* act, const float* W, int M, int K,
             float* dX) {
  hipLaunchKernelGGL(k_gemm_dx<NOUT>, dim3(blocks(M, 64)), dim3(256), 0, S->stream, dY, act, W, M,
                     K, dX);
}

void forward(storm_mpnn* S) {
  const int N = S->N, E = S->E, L = S->cfg.layers, Ke = S->Ke();
  const float* P = S->p.p;
  gemm<HID>(S, S->xin.p, P + S->o_enc_w, P + S->o_enc_b, N, S->cfg.f_node, 1, S->H[0].p);
  for (int l = 0; l < L; ++l) {
    hipLaunchKernelGGL(k_edge_in, dim3(blocks((size_t)E * Ke)), dim3(256), 0, S->stream, S->H[l].p,
                       S->ef.p, S->src.p, S->dst.p, E, S->cfg.f_edge, S->Xe[l].p);
    gemm<HID>(S, S->Xe[l].p, P + S->o_ew[l], P + S->o_eb[l], E, Ke, 1, S->Mg[l].p);
    hipLaunchKernelGGL(k_node_in, dim3(blocks((size_t)N * HID)), dim3(256), 0, S->stream,
                       S->H[l].p, S->Mg[l].p, S->in_ptr.p, S->in_idx.p, N, S->Xn[l].p);
    gemm<HID>(S, S->Xn[l].p, P + S->o_nw[l], P + S->o_nb[l], N, 2 * HID, 1, S->U[l].p);
    hipLaunchKernelGGL(k_add, dim3(blocks((size_t)N * HID)), dim3(256), 0, S->stream, S->H[l].p,
                       S->U[l].p, (size_t)N * HID, S->H[l + 1].p);
  }
  if (S->cfg.f_out == 16)
    gemm<16>(S, S->H[L].p, P + S->o_dec_w, P + S->o_dec_b, N, HID, 0, S->Y.p);
  else
    throw Err{"f_out must be 16"};
  HIPCHK(hipGetLastError());
}

void backward(storm_mpnn* S) {
  const int N = S->N, E = S->E, L = S->cfg.layers, Ke = S->Ke();
  const float* P = S->p.p;
  float* G = S->g.p;
  // decoder
  weight_grad<16>(S, S->H[L].p, S->dY.p, nullptr, N, HID, G + S->o_dec_w, G + S->o_dec_b);
  gemm_dx<16>(S, S->dY.p, nullptr, P + S->o_dec_w, N, HID, S->dH.p);  // dH^L
  for (int l = L - 1; l >= 0; --l) {
    // h^{l+1} = h^l + u: du = dh^{l+1} masked by u > 0
    weight_grad<HID>(S, S->Xn[l].p, S->dH.p, S->U[l].p, N, 2 * HID, G + S->o_nw[l], G + S->o_nb[l]);
    gemm_dx<HID>(S, S->dH.p, S->U[l].p, P + S->o_nw[l], N, 2 * HID, S->dXn.p);
    hipLaunchKernelGGL(k_edge_grad, dim3(blocks((size_t)E * HID)), dim3(256), 0, S->stream,
                       S->dXn.p, S->dst.p, E, S->dM.p);
    weight_grad<HID>(S, S->Xe[l].p, S->dM.p, S->Mg[l].p, E, Ke, G + S->o_ew[l], G + S->o_eb[l]);
    gemm_dx<HID>(S, S->dM.p, S->Mg[l].p, P + S->o_ew[l], E, Ke, S->dXe.p);
    hipLaunchKernelGGL(k_node_grad, dim3(blocks((size_t)N * HID)), dim3(256), 0, S->stream,
                       S->dH.p, S->dXn.p, S->dXe.p, Ke, S->out_ptr.p, S->out_idx.p, S->in_ptr.p,
                       S->in_idx.p, N, S->dH2.p);
    std::swap(S->dH.p, S->dH2.p);
    std::swap(S->dH.n, S->dH2.n);
  }
  // encoder: h0 = relu(x W + b)
  weight_grad<HID>(S, S->xin.p, S->dH.p, S->H[0].p, N, S->cfg.f_node, G + S->o_enc_w,
                   G + S->o_enc_b);
  HIPCHK(hipGetLastError());
}

int fail(storm_mpnn* S, const std::string& m) {
  if (S) S->err = m;
  return -1;
}

}  // namespace

extern "C" {

int storm_mpnn_create(const storm_mpnn_cfg_t* cfg, storm_mpnn_t** out) {
  if (!cfg || !out) return -1;
  *out = nullptr;
  if (cfg->hidden != HID || cfg->layers < 1 || cfg->layers > 16 || cfg->f_node % 16 ||
      cfg->f_edge % 16 || cfg->f_node <= 0 || cfg->f_edge <= 0 || cfg->f_out != 16) {
    std::fprintf(stderr, "storm_mpnn_create: hidden must be 64, f_node/f_edge multiples of 16, "
                         "f_out 16, 1..16 layers\n");
    return -1;
  }
  std::unique_ptr<storm_mpnn> S(new (std::nothrow) storm_mpnn());
  if (!S) return -1;
  try {
    S->cfg = *cfg;
    HIPCHK(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
    const int L = cfg->layers, Ke = S->Ke();
    int64_t o = 0;
    S->o_enc_w = o; o += (int64_t)cfg->f_node * HID;
    S->o_enc_b = o; o += HID;
    for (int l = 0; l < L; ++l) {
      S->o_ew.push_back(o); o += (int64_t)Ke * HID;
      S->o_eb.push_back(o); o += HID;
      S->o_nw.push_back(o); o += (int64_t)2 * HID * HID;
      S->o_nb.push_back(o); o += HID;
    }
    S->o_dec_w = o; o += (int64_t)HID * cfg->f_out;
    S->o_dec_b = o; o += cfg->f_out;
    S->np = o;
    // He-uniform weights, zero biases, zero decoder
    std::vector<float> h(o, 0.f);
    uint64_t st = cfg->seed;
    auto fill = [&](int64_t off, int fan_in, int fan_out) {
      const double a = std::sqrt(6.0 / fan_in);
      for (int64_t i = 0; i < (int64_t)fan_in * fan_out; ++i) {
        const double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
        h[off + i] = (float)((2.0 * u - 1.0) * a);
      }
    };
    fill(S->o_enc_w, cfg->f_node, HID);
    for (int l = 0; l < L; ++l) {
      fill(S->o_ew[l], Ke, HID);
      fill(S->o_nw[l], 2 * HID, HID);
    }
    // the decoder starts at zero: an untrained network outputs zero
    S->p.need(o); S->g.need(o); S->m.need(o); S->v.need(o);
    HIPCHK(hipMemcpy(S->p.p, h.data(), sizeof(float) * o, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(S->g.p, 0, sizeof(float) * o));
    HIPCHK(hipMemset(S->m.p, 0, sizeof(float) * o));
    HIPCHK(hipMemset(S->v.p, 0, sizeof(float) * o));
    S->H.resize(L + 1); S->Xe.resize(L); S->Mg.resize(L); S->Xn.resize(L); S->U.resize(L);
  } catch (const Err& e) {
    std::fprintf(stderr, "storm_mpnn_create: %s\n", e.msg.c_str());
    return -1;
  }
  *out = S.release();
  return 0;
}

void storm_mpnn_destroy(storm_mpnn_t* S) { delete S; }

int64_t storm_mpnn_num_params(const storm_mpnn_t* S) { return S ? S->np : -1; }

int storm_mpnn_get_params(storm_mpnn_t* S, float* p) {
  if (!S || !p) return -1;
  try {
    HIPCHK(hipStreamSynchronize(S->stream));
    HIPCHK(hipMemcpy(p, S->p.p, sizeof(float) * S->np, hipMemcpyDeviceToHost));
  } catch (const Err& e) {
    return fail(S, e.msg);
  }
  return 0;
}

int storm_mpnn_set_params(storm_mpnn_t* S, const float* p) {
  if (!S || !p) return -1;
  try {
    HIPCHK(hipStreamSynchronize(S->stream));
    HIPCHK(hipMemcpy(S->p.p, p, sizeof(float) * S->np, hipMemcpyHostToDevice));
  } catch (const Err& e) {
    return fail(S, e.msg);
  }
  return 0;
}

int storm_mpnn_set_graph(storm_mpnn_t* S, int32_t n, int32_t e, const int32_t* src,
                         const int32_t* dst, const float* nf, const float* efe) {
  if (!S || n < 1 || e < 1 || !src || !dst || !efe || !nf) return -1;
  try {
    for (int k = 0; k < e; ++k)
      if (src[k] < 0 || src[k] >= n || dst[k] < 0 || dst[k] >= n)
        return fail(S, "edge endpoint out of range");
    HIPCHK(hipStreamSynchronize(S->stream));
    S->N = n;
    S->E = e;
    // CSR of the edges into / out of every node, ascending edge ids
    std::vector<int> ip(n + 1, 0), op(n + 1, 0), ii(e), oi(e);
    for (int k = 0; k < e; ++k) { ip[dst[k] + 1]++; op[src[k] + 1]++; }
    for (int i = 0; i < n; ++i) { ip[i + 1] += ip[i]; op[i + 1] += op[i]; }
    {
      std::vector<int> fi(ip.begin(), ip.end() - 1), fo(op.begin(), op.end() - 1);
      for (int k = 0; k < e; ++k) { ii[fi[dst[k]]++] = k; oi[fo[src[k]]++] = k; }
    }
    auto up = [&](DBuf<int>& d, const int* h, size_t k) {
      d.need(k);
      if (k) HIPCHK(hipMemcpy(d.p, h, sizeof(int) * k, hipMemcpyHostToDevice));
    };
    up(S->src, src, e); up(S->dst, dst, e);
    up(S->in_ptr, ip.data(), n + 1); up(S->in_idx, ii.data(), e);
    up(S->out_ptr, op.data(), n + 1); up(S->out_idx, oi.data(), e);
    const int fn = S->cfg.f_node, fe = S->cfg.f_edge, Ke = S->Ke(), L = S->cfg.layers;
    S->xin.need((size_t)n * fn);
    HIPCHK(hipMemcpy(S->xin.p, nf, sizeof(float) * n * fn, hipMemcpyHostToDevice));
    S->ef.need(std::max<size_t>((size_t)e * fe, 1));
    if (e) HIPCHK(hipMemcpy(S->ef.p, efe, sizeof(float) * e * fe, hipMemcpyHostToDevice));
    for (int l = 0; l <= L; ++l) S->H[l].need((size_t)n * HID);
    for (int l = 0; l < L; ++l) {
      S->Xe[l].need((size_t)e * Ke + 1);
      S->Mg[l].need((size_t)e * HID + 1);
      S->Xn[l].need((size_t)n * 2 * HID);
      S->U[l].need((size_t)n * HID);
    }
    S->Y.need((size_t)n * S->cfg.f_out);
    S->dY.need((size_t)n * S->cfg.f_out);
    S->dH.need((size_t)n * HID); S->dH2.need((size_t)n * HID);
    S->dXn.need((size_t)n * 2 * HID);
    S->dM.need((size_t)e * HID + 1);
    S->dXe.need((size_t)e * Ke + 1);
    S->have_fwd = false;
  } catch (const Err& er) {
    return fail(S, er.msg);
  }
  return 0;
}

int storm_mpnn_forward(storm_mpnn_t* S, float* out) {
  if (!S || S->N < 1) return -1;
  try {
    forward(S);
    S->have_fwd = true;
    if (out) {
      HIPCHK(hipMemcpyAsync(out, S->Y.p, sizeof(float) * S->N * S->cfg.f_out, hipMemcpyDeviceToHost,
                            S->stream));
    }
    HIPCHK(hipStreamSynchronize(S->stream));
  } catch (const Err& e) {
    return fail(S, e.msg);
  }
  return 0;
}

int storm_mpnn_backward(storm_mpnn_t* S, const float* target, const float* w, double* loss) {
  if (!S || !target || !S->have_fwd) return fail(S, "backward needs a forward pass and a target");
  try {
    const int N = S->N, F = S->cfg.f_out;
    DBuf<float>& T = S->tgt;
    DBuf<float>& Wt = S->wgt;
    T.need((size_t)N * F);
    HIPCHK(hipMemcpyAsync(T.p, target, sizeof(float) * N * F, hipMemcpyHostToDevice, S->stream));
    if (w) {
      Wt.need(N);
      HIPCHK(hipMemcpyAsync(Wt.p, w, sizeof(float) * N, hipMemcpyHostToDevice, S->stream));
    }
    const unsigned nb = blocks((size_t)N * F);
    S->lossp.need(nb);
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, S->stream, S->Y.p, T.p, w ? Wt.p : nullptr, N,
                       F, S->dY.p, S->lossp.p);
    backward(S);
    std::vector<double> lp(nb);
    HIPCHK(hipMemcpyAsync(lp.data(), S->lossp.p, sizeof(double) * nb, hipMemcpyDeviceToHost,
                          S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    double s = 0.0;
    for (double x : lp) s += x;
    if (loss) *loss = s / ((double)N * F);
  } catch (const Err& e) {
    return fail(S, e.msg);
  }
  return 0;
}

int storm_mpnn_get_grads(storm_mpnn_t* S, float* g) {
  if (!S || !g) return -1;
  try {
    HIPCHK(hipStreamSynchronize(S->stream));
    HIPCHK(hipMemcpy(g, S->g.p, sizeof(float) * S->np, hipMemcpyDeviceToHost));
  } catch (const Err& e) {
    return fail(S, e.msg);
  }
  return 0;
}

int storm_mpnn_adamw_step(storm_mpnn_t* S) {
  if (!S) return -1;
  try {
    S->step++;
    const storm_mpnn_cfg_t& c = S->cfg;
    const float c1 = 1.f - std::pow(c.beta1, (float)S->step);
    const float c2 = 1.f - std::pow(c.beta2, (float)S->step);
    hipLaunchKernelGGL(k_adamw, dim3(blocks(S->np)), dim3(256), 0, S->stream, S->p.p, S->g.p,
                       S->m.p, S->v.p, S->np, c.lr, c.beta1, c.beta2, c.eps, c.weight_decay, c1, c2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(S->stream));
  } catch (const Err& e) {
    return fail(S, e.msg);
  }
  return 0;
}

int storm_mpnn_set_lr(storm_mpnn_t* S, float lr) {
  if (!S || !(lr >= 0.f)) return -1;
  S->cfg.lr = lr;
  return 0;
}

const char* storm_mpnn_last_error(storm_mpnn_t* S) { return S ? S->err.c_str() : "null network"; }

int storm_mpnn_gvf_profile(double q, double S, double manning, double h0, double h_floor,
                           double ds, int32_t n_steps, double* h) {
  if (!h || n_steps < 0) return -1;
  constexpr double G = 9.81;
  const double third10 = 10.0 / 3.0;
  const auto rate = [&](double y) {
    if (h_floor > y) y = h_floor;
    const double fr2 = q * q / (G * std::pow(y, 3.0));
    const double sf = manning * manning * q * q / std::pow(y, third10);
    return -(S - sf) / (1.0 - fr2);
  };
  h[0] = h0;
  for (int32_t k = 0; k < n_steps; ++k) {
    const double y = h[k];
    const double k1 = rate(y);
    const double k2 = rate(y + 0.5 * ds * k1);
    const double k3 = rate(y + 0.5 * ds * k2);
    const double k4 = rate(y + ds * k3);
    const double y1 = y + ds * (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0;
    h[k + 1] = (h_floor > y1) ? h_floor : y1;
  }
  return 0;
}

}  // extern "C"
