// MI355X (gfx950) FP64 kernels of the SToRM FVT time step and the C ABI
// declared in include/storm_fvt.h.
//
// Per RK stage (solverFVT.f90:158-331) two launches:
//   k_vertex_stage (k_vertex_first in the first stage)
//             the IDW water surface, depth and wet flag of every vertex
//             (to_vertices.f90:26-41 with the stage strings' nodes of
//             enforce_bcs);
//   k_stage   one block per tile (a run of Hilbert-ordered cells and its
//             halo of edge neighbours):
//   - find_dry_cells' flags, cell_grads (LSQ/Gauss + Barth-Jespersen),
//     hgradfix2 and to_vertices' cell pass for the tile and halo cells, whose
//     face states stay in shared memory;
//   - inflowQbyH's fixed point for the strings the tile touches;
//   - every flux of each tile cell's three sides, gathered by the cell in the
//     reference's accumulation order (no atomics), bed slope, explicit
//     friction if chosen, SSP-RK update, clip and wet/dry recovery; in the
//     last stage also implicit friction, hvalve, the residual partial sums
//     and the next step's CFL candidate.
// Per step the bookkeeping (dt, time, BC tables, stopping; one block per
// reach) runs as the first blocks of the first stage's vertex launch, and a
// standalone k_control closes each run.  The Green-Gauss option adds its
// first pass (k_gauss1) before each stage.
//
// Per-cell records are 32-byte structs (static, stage state); per-cell edge
// data are stored per (slot, cell).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/storm_fvt.h"
#include "fvt_math.h"
#include "fvt_prep.h"

using namespace storm;

namespace {

constexpr int CTRL_THREADS = 1024;
constexpr int DEFAULT_CAP = 256;  // cells + halo cells per tile (threads per stage block)
// Waves per SIMD the stage kernel is compiled for: four at capacity 256 (four
// blocks per CU), the compiler's choice otherwise (the xsf kernel's shared
// memory and the larger capacities allow fewer blocks per CU anyway).
#define STAGE_WAVES(cap, xsf) ((cap) == 256 && !(xsf) ? 4 : 1)

// largest power of two not above n (the fixed tree shape of block reductions)
__host__ __device__ constexpr int pow2_floor(int n) {
  int p = 1;
  while (2 * p <= n) p *= 2;
  return p;
}

// Per-tile partials of the last stage, PART doubles per tile: sum (h - h^n)^2,
// sum (h^n)^2, max|phi|, the next step's CFL candidate, and (stopping test 2
// only) sum |q - q^n|^2 and sum |q^n|^2 of the discharge q = (hu, hv).
constexpr int PART = 6;

// Step control of one reach, kept on the device (k_control writes it).
struct Ctrl {
  double dt;          // dt of the current step
  double delt_last;   // steady: local dt of the reference's last cell
  double ttime, stime, rmax, sumtop, sumbot;
  int iter, done, pending, pad;
};

// Per reach: its strings, its tiles, the cell whose local dt drives steady
// friction, and bch_max (dataprepFVT.f90:278-286).
struct RInfo {
  int in_base, n_in, out_base, n_out, tile0, tile1, last_cell, pad;
  double bch_max, pad2;
};

// Cell records, read whole with 16-byte loads: static data and the stage
// state (double-buffered between stages).
struct alignas(32) CellS { double xc, yc, z, cd; };
struct alignas(32) CellD { double h, u, v, dry; };  // dry: 1 or 0

struct Geo {
  int N, P, R;
  const int *cvtx, *cnbr;
  const uint64_t* cmeta;
  const CellS* cs;
  const double4* vxyz;
  const int *mptr, *mcell;
  const int* gconn;
  const double *gcx, *gcy, *gwt;
  const int *vptr, *vcell, *vbc;
  const int* vell;          // [vW][P] incident cells per vertex, n2t order
  const double* vwt;        // [vW][P] inverse-distance weights
  const double2* vzb;       // per vertex {zvtx, bits: string index | degree << 32}
  // per chunk of 256 vertices (one block of the vertex pass), the tiles that
  // own its vertices' incident cells: vc_tile[vc_ptr[c] .. vc_ptr[c+1])
  const int *vc_ptr, *vc_tile;
  const int *qs_ptr, *qs_cell, *qs_slot;
  const double *qs_rcx, *qs_rcy, *qs_nx, *qs_ny, *qs_len;
  const int* in_type;
  const int *svptr, *svtx;  // per reach, the vertices of its stage strings
  const RInfo* reach;
  int n_in, n_out, nr;  // string totals, reaches
  const double2* wind;  // per cell wcoef1, wcoef2 (windforcing.f90), when Opts::wind
  const int2* cv_io;    // per culvert its inlet and outlet cells (internal ids)
  const double* cvsrc;  // per culvert the discharge of the current stage
};

struct Opts {
  int rkorder, dt_op, steady, stop_cr, btype, hvalve, limiter, flux, gradient, friction, vclip;
  int nonfinite_stop;
  int xsf;  // interior edges as invisc_fluxes_xsf.f90 forms them
  int wind;  // wind terms on (Geo::wind)
  int ncv;   // culverts (Geo::cv_io, cvsrc)
  double cfl, rstop, h_dry, h_wet, vceiling;
};

struct Stage {
  const double* slot[4];  // rk slots 0..s-1 (h at +0, hu at +N, hv at +2N)
  double alpha[4];
  double* out;            // slot s
  double* hin;            // slot s-1 (its hu, hv are zeroed by hgradfix2)
  double beta;
  int s, last;
};

struct State {
  const CellD* din;   // state at the start of the stage
  CellD* dout;        // state for the next stage
  double* zeta;       // water surface of the next stage's input, NaN at dry cells
  uint8_t* twet;      // per tile: some owned cell's zeta is not NaN (written with zeta)
  double *gwork, *hvtx, *zetavtx, *phi, *part;
  uint8_t *dry_post, *vwet;
  double *sq, *sh;    // current Q (or stage) of every inflow string, stage of every outflow string
  double *phi0, *phis;  // residual smoothing: the residual before and after it
  // the shared-memory sweep's result in level order and each cell's
  // position in it (null: the result is phis, in cell order)
  const double4* phis_lo;
  const int* rs_pos;
  Ctrl* ctrl;         // one per reach
  // per reach, `done` as the first stage of the latest step saw it (that
  // stage's kernel writes it), for the next step's first vertex pass, which
  // runs while k_control may be writing `done`
  int* vdone;
};

// Tiles of the stage kernel: tile b owns cells [tptr[b], tptr[b+1]) (a run of
// the Hilbert-ordered cells) and its halo of edge neighbours outside the run
// is hcell[hptr[b]..hptr[b+1]).  Local cell ids are 16 bits, three to a word
// (cell slot j at bit 16*j): tile cells 0..T-1, then halo cells.
struct Tiles {
  const int4* hdr;              // per block, two int4: {c0, T, hb, H}, {qmask lo, hi, tile, 0}
  const int4* hrec;             // per halo entry, two int4: {cell, nbr0..2}, {vtx0..2, 0}
  const uint64_t* hmeta;        // per halo entry: the cell's metadata word
  const uint64_t* cln;          // per tile cell: local neighbour ids (0xFFFF: none)
  // per block {first, interior count | boundary count << 16} of its side
  // records (the sides whose fluxes are formed once each: interior edges with
  // an owned cell, then owned boundary sides); record bits 0-15 owned local id a, 16-17 side ja, 18-35 the slot
  // bits of (a, ja), 36-51 the neighbour's local id across an interior side
  const int2* ehdr;
  const uint64_t* erec;
  int nt;
};

__device__ __forceinline__ CellD ld(const CellD* p) {
  const double4 q = *reinterpret_cast<const double4*>(p);
  return CellD{q.x, q.y, q.z, q.w};
}
__device__ __forceinline__ CellS ld(const CellS* p) {
  const double4 q = *reinterpret_cast<const double4*>(p);
  return CellS{q.x, q.y, q.z, q.w};
}
__device__ __forceinline__ void st4(CellD* p, double a, double b, double c, double d) {
  *reinterpret_cast<double4*>(p) = make_double4(a, b, c, d);
}

// ---------------------------------------------------------------- XCD-aware tiles
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Mapping dispatch index b to tile start(b % 8) + b / 8 gives every XCD one
// contiguous range of the (RCM-ordered) cells, so the neighbour gathers of a
// tile hit the L2 that fetched the neighbouring tiles.  The map is a
// bijection on [0, gridDim.x).
constexpr int NXCD = 8;
__device__ __forceinline__ int xcd_map(int nb, int b) {
  const int q = nb / NXCD, r = nb % NXCD;
  const int x = b % NXCD, k = b / NXCD;
  return x * q + (x < r ? x : r) + k;
}
__device__ __forceinline__ int xcd_tile() { return xcd_map(gridDim.x, blockIdx.x); }

// ---------------------------------------------------------------- cell geometry
// Everything geometric is formed on the fly from the three vertices, with the
// operations and operand order of the reference set-up routine cited, so the
// values equal the ones the reference stores.
struct Tri {
  double x[3], y[3], zv[3];
};

__device__ __forceinline__ Tri load_tri(const Geo& g, int i) {
  Tri t;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double4 q = g.vxyz[g.cvtx[k * g.N + i]];
    t.x[k] = q.x; t.y[k] = q.y; t.zv[k] = q.z;
  }
  return t;
}

// t_areas.f90:40-41
__device__ __forceinline__ double tri_area(const Tri& t) {
  return HALF * fabs((t.x[0] - t.x[1]) * (t.y[0] + t.y[1]) + (t.x[1] - t.x[2]) * (t.y[1] + t.y[2]) +
                     (t.x[2] - t.x[0]) * (t.y[2] + t.y[0]));
}

// find_edges.f90:159-163 (length), edge_normals.f90:35-38 + dataprepFVT.f90:176-181
// (unit normal and tangent of the side from vertex j to vertex j+1; the
// reference's edges(k) normal is this one, negated when edges(k)%p runs
// the other way, and flipped to e(1)'s orientation before use)
struct Side {
  double len, nx, ny, tx, ty, mx, my, zm;
};
__device__ __forceinline__ Side side_of(double xa, double ya, double za, double xb, double yb,
                                       double zb) {
  Side s;
  const double dx = xa - xb, dy = ya - yb;
  s.len = sqrt(dx * dx + dy * dy);
  const Recip rl = recip(s.len, yb - ya);
  s.nx = qdiv(rl, yb - ya, s.len);
  s.ny = qdiv(rl, xa - xb, s.len);
  s.tx = -s.ny;  // (xb - xa) / len, the same value
  s.ty = s.nx;   // (yb - ya) / len
  s.mx = HALF * (xa + xb);  // dataprepFVT.f90:168-169
  s.my = HALF * (ya + yb);
  s.zm = HALF * (za + zb);  // invisc_fluxes.f90:89
  return s;
}
template <int J>
__device__ __forceinline__ Side side(const Tri& t) {
  constexpr int A = J, B = (J + 1) % 3;
  return side_of(t.x[A], t.y[A], t.zv[A], t.x[B], t.y[B], t.zv[B]);
}


__device__ __forceinline__ uint32_t slot_bits(uint64_t meta, int j) {
  return (uint32_t)(meta >> (CM_SLOT0 + CM_SLOT_BITS * j)) & ((1u << CM_SLOT_BITS) - 1);
}

__device__ __forceinline__ bool find_dry(double h, double u, double v, double h_dry, double h_wet,
                                         double* umag) {
  // find_dry_cells.f90:25-42
  bool d = h < h_dry;
  double um = ZERO;
  if (!d) um = sqrt(u * u + v * v);
  if (!d && h < h_wet && equals(um, ZERO)) d = true;
  *umag = um;
  return d;
}

// ---------------------------------------------------------------- init
template <int CAP>
__global__ void __launch_bounds__(CAP)
k_init(Geo g, Opts o, Tiles tl, double* slot0, const double* h0, const double* u0,
       const double* v0, CellD* d, double* zeta, uint8_t* twet, double* part) {
  // solverFVT.f90:123-127 and the first find_dry_cells, one block per tile;
  // the CFL candidates of step 1 go to the per-tile partials (slot 3),
  // reduced by k_control
  __shared__ double s_c[CAP];
  const int t = threadIdx.x;
  const int4 hd0 = tl.hdr[2 * blockIdx.x], hd1 = tl.hdr[2 * blockIdx.x + 1];
  const int c0 = hd0.x, T = hd0.y, b = hd1.z;
  double cand = DBL_MAX;
  bool wet_out = false;
  if (t < T) {
    const int i = c0 + t;
    const double h = h0[i], uu = u0[i], vv = v0[i];
    slot0[g.N + i] = h * uu;
    slot0[2 * g.N + i] = h * vv;
    slot0[i] = h;
    double um;
    const bool dr = find_dry(h, uu, vv, o.h_dry, o.h_wet, &um);
    st4(&d[i], h, uu, vv, dr ? ONE : ZERO);
    zeta[i] = dr ? __builtin_nan("") : h + g.cs[i].z;
    wet_out = !dr;
    const double wc = dr ? ZERO : sqrt(G * h);
    const Tri tr = load_tri(g, i);
    const double area = tri_area(tr);
    const double perim = side<0>(tr).len + side<1>(tr).len + side<2>(tr).len;
    const double dd = um + wc + SMALL;
    cand = o.cfl * area / (dd * perim);
  }
  const int any_wet = __syncthreads_or(wet_out);
  if (t == 0) twet[b] = any_wet != 0;
  s_c[t] = cand;
  __syncthreads();
  constexpr int P2 = pow2_floor(CAP);
  if (P2 < CAP) {
    if (t < CAP - P2) s_c[t] = dmin(s_c[t], s_c[t + P2]);
    __syncthreads();
  }
  for (int w = P2 / 2; w > 0; w >>= 1) {
    if (t < w) s_c[t] = dmin(s_c[t], s_c[t + w]);
    __syncthreads();
  }
  if (t == 0) {
    part[PART * b] = ZERO;
    part[PART * b + 1] = ZERO;
    part[PART * b + 2] = ZERO;
    part[PART * b + 3] = s_c[0];
    part[PART * b + 4] = ZERO;
    part[PART * b + 5] = ZERO;
  }
}

// ---------------------------------------------------------------- control
__device__ void get_time_bcs(const RInfo& ri, const int* out_type, const int* qts_ptr,
                             const double* qts_t, const double* qts_v, const int* hts_ptr,
                             const double* hts_t, const double* hts_v, double time, double* sq,
                             double* sh) {
  // get_time_bcs.f90:31-77 for the strings of one reach
  for (int k = ri.out_base; k < ri.out_base + ri.n_out; ++k) {
    if (out_type[k] != 1) continue;
    int a = hts_ptr[k], n = hts_ptr[k + 1] - a;
    if (n <= 0) continue;
    const double* tl = hts_t + a;
    const double* tv = hts_v + a;
    if (time < tl[0]) sh[k] = tv[0];
    else if (time >= tl[n - 1]) sh[k] = tv[n - 1];
    else
      for (int i = 0; i < n - 1; ++i)
        if (time >= tl[i] && time < tl[i + 1]) {
          double d = time - tl[i], d2 = tl[i + 1] - tl[i];
          sh[k] = tv[i] + d * (tv[i + 1] - tv[i]) / d2;
          break;
        }
  }
  for (int k = ri.in_base; k < ri.in_base + ri.n_in; ++k) {
    int a = qts_ptr[k], n = qts_ptr[k + 1] - a;
    if (n <= 0) continue;
    const double* tl = qts_t + a;
    const double* tv = qts_v + a;
    if (time < tl[0]) sq[k] = tv[0];
    else if (time >= tl[n - 1]) sq[k] = tv[n - 1];
    else
      for (int i = 0; i < n - 1; ++i)
        if (time >= tl[i] && time < tl[i + 1]) {
          double d = time - tl[i], d2 = tl[i + 1] - tl[i];
          sq[k] = tv[i] + d * (tv[i + 1] - tv[i]) / d2;
          break;
        }
  }
}

// What k_control reads and writes besides the mesh and options.
struct CtrlArgs {
  Ctrl* ctrl;
  const double* part;
  double *sq, *sh, *hvtx;
  int begin;
  const int *out_type, *qts_ptr, *hts_ptr;
  const double *qts_t, *qts_v, *hts_t, *hts_v;
};

// Step bookkeeping of solverFVT.f90:132-155 (begin) and :338-472 (end of the
// previous step: residuals, stopping test, time, boundary tables) for reach
// r, by a block of NT threads.
template <int NT>
__device__ void control_block(const Geo& g, const Opts& o, const CtrlArgs& A, int r) {
  const RInfo ri = g.reach[r];
  Ctrl* c = &A.ctrl[r];
  const double* part = A.part;
  double* sh = A.sh;
  constexpr int NW = NT / 64;
  __shared__ double s_top[NW], s_bot[NW], s_max[NW], s_dt[NW], s_topq[NW], s_botq[NW];
  __shared__ double s_str[MAX_STRINGS];
  __shared__ int s_go;
  const int t = threadIdx.x;
  const Ctrl cq = *c;  // one load of the whole control block
  if (cq.done) return;
  // fixed-order reduction of the per-block partials of the last stage
  // (residual sums, max|phi|) and of the next step's CFL candidates
  double top = ZERO, bot = ZERO, mx = ZERO, dtm = DBL_MAX, topq = ZERO, botq = ZERO;
  // tiles t, t + NT, ... in that order.  UR tiles per round are loaded
  // unconditionally (an index past the reach's last tile reads the last tile
  // again and is then dropped), so a round's loads are in flight together;
  // UR is small because the first vertex pass shares this kernel's register
  // budget
  constexpr int UR = 4;
  const int last = ri.tile1 - 1;
  for (int b0 = ri.tile0 + t; b0 <= last; b0 += UR * NT) {
    double2 v[UR][2];
#pragma unroll
    for (int u = 0; u < UR; ++u) {
      const double2* q = reinterpret_cast<const double2*>(&part[PART * min(b0 + u * NT, last)]);
      v[u][0] = q[0];
      v[u][1] = q[1];
    }
#pragma unroll
    for (int u = 0; u < UR; ++u) {
      const bool in = b0 + u * NT <= last;
      top = top + (in ? v[u][0].x : ZERO);
      bot = bot + (in ? v[u][0].y : ZERO);
      mx = in ? dmax(mx, v[u][1].x) : mx;
      dtm = in ? dmin(dtm, v[u][1].y) : dtm;
    }
  }
  if (o.stop_cr == 2)
    for (int b = ri.tile0 + t; b <= last; b += NT) {
      topq = topq + part[PART * b + 4];
      botq = botq + part[PART * b + 5];
    }
  // a fixed butterfly within each wave (every lane ends with the same
  // bits), then the waves' results in wave order
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    top = top + __shfl_xor(top, off);
    bot = bot + __shfl_xor(bot, off);
    mx = dmax(mx, __shfl_xor(mx, off));
    dtm = dmin(dtm, __shfl_xor(dtm, off));
    topq = topq + __shfl_xor(topq, off);
    botq = botq + __shfl_xor(botq, off);
  }
  if ((t & 63) == 0) {
    const int w = t >> 6;
    s_top[w] = top; s_bot[w] = bot; s_max[w] = mx; s_dt[w] = dtm;
    s_topq[w] = topq; s_botq[w] = botq;
  }
  __syncthreads();
  if (t == 0) {
#pragma unroll
    for (int w = 1; w < NW; ++w) {
      s_top[0] = s_top[0] + s_top[w];
      s_bot[0] = s_bot[0] + s_bot[w];
      s_max[0] = dmax(s_max[0], s_max[w]);
      s_dt[0] = dmin(s_dt[0], s_dt[w]);
      s_topq[0] = s_topq[0] + s_topq[w];
      s_botq[0] = s_botq[0] + s_botq[w];
    }
    Ctrl cc = cq;  // thread 0 works on a register copy, stored back once
    bool go = true;
    if (cc.pending) {
      double sumtop = s_top[0], sumbot = s_bot[0];
      cc.sumtop = sumtop;
      cc.sumbot = sumbot;
      if (o.steady) {
        if (o.stop_cr == 0) cc.rmax = s_max[0];
        else {
          if (equals(sumbot, ZERO)) sumbot = ONE;
          cc.rmax = sqrt(sumtop / sumbot);
          if (o.stop_cr == 2) {
            // the same relative RMS change of the discharge (no reference
            // counterpart: stopping test 1 looks at h alone)
            double sumbotq = s_botq[0];
            if (equals(sumbotq, ZERO)) sumbotq = ONE;
            cc.rmax = dmax(cc.rmax, sqrt(s_topq[0] / sumbotq));
          }
        }
        cc.stime = (double)cc.iter;
        cc.pending = 0;
        if (cc.rmax < o.rstop) { cc.done = 1; go = false; }
        else if (o.nonfinite_stop && !(cc.rmax <= DBL_MAX)) {  // NaN or infinite: cannot converge
          cc.rmax = __builtin_nan("");
          cc.done = 1;
          go = false;
        }
      } else {
        cc.stime = cc.stime + cc.dt;
        cc.rmax = s_max[0];
        get_time_bcs(ri, A.out_type, A.qts_ptr, A.qts_t, A.qts_v, A.hts_ptr, A.hts_t, A.hts_v,
                     cc.stime, A.sq, sh);
        cc.pending = 0;
      }
    }
    if (go && A.begin) {
      cc.iter = cc.iter + 1;
      if (!o.steady) {
        cc.dt = (o.dt_op == 0) ? o.cfl : s_dt[0];
        cc.ttime = cc.ttime + cc.dt;
      } else {
        cc.ttime = cc.ttime + cc.delt_last;
      }
      cc.pending = 1;
    }
    *c = cc;
    s_go = go;
    // the stage of every stage string at the new time, for the pass below
    for (int k = 0; k < ri.n_out; ++k) s_str[k] = sh[ri.out_base + k];
  }
  __syncthreads();
  if (!s_go) return;
  // enforce_bcs.f90:40-56 for the nodes of the reach's stage strings, at the
  // step's new time: the step's first vertex pass runs beside this block and
  // leaves these vertices' depths to it
  for (int k = g.svptr[r] + t; k < g.svptr[r + 1]; k += NT) {
    const int p = g.svtx[k];
    const double zv = g.vxyz[p].z;
    A.hvtx[p] = dmax(ZERO, s_str[g.vbc[p] - ri.out_base] - zv);
  }
}

// k_control alone, one block per reach (the end of a run)
__global__ void __launch_bounds__(CTRL_THREADS) k_control(Geo g, Opts o, CtrlArgs A) {
  control_block<CTRL_THREADS>(g, o, A, blockIdx.x);
}

// ---------------------------------------------------------------- vertex pass
// to_vertices.f90:26-41 with ctr2vtxWSE.f90:89-105 (inverse-distance weights
// of c2v_setup.f90:106-113) and positivity_vtx.f90 for vertex p, then the
// stage strings' nodes of enforce_bcs.f90:40-56.  `dry` overrides the dry
// flags of `d` when given (the post-loop call uses the flags of the
// post-step find_dry_cells).
struct Vtx {
  double h, zeta;
  bool wet;
};

__device__ __forceinline__ Vtx vertex_idw(const Geo& g, const Opts& o, const CellD* d,
                                          const uint8_t* dry, int p, const double4& q,
                                          const double* sh) {
  double id = ZERO, a = ZERO, zr = q.z;
  int counter = 0;
  const int t0 = g.vptr[p], t1 = g.vptr[p + 1];
  for (int t = t0; t < t1; t += 4) {
    // four cells per round, all loads issued before the sums
    int k[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) k[u] = (t + u < t1) ? g.vcell[t + u] : -1;
    CellD dk[4];
    CellS sk[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int kk = k[u] < 0 ? k[0] : k[u];
      dk[u] = ld(&d[kk]);
      sk[u] = ld(&g.cs[kk]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (k[u] < 0) continue;
      const bool kd = dry ? (dry[k[u]] != 0) : (dk[u].dry != ZERO);
      if (kd) continue;
      counter++;
      const double xk = sk[u].xc - q.x, yk = sk[u].yc - q.y;
      const double w = ONE / sqrt(xk * xk + yk * yk);
      const double zk = dk[u].h + sk[u].z;
      id = id + w;
      a = a + w * zk;
      zr = dmax(zk, zr);
    }
  }
  double zeta = (counter > 0) ? a / id : q.z;
  zeta = dmin(zeta, zr);
  double hv = zeta - q.z;
  if (hv < ZERO) { hv = ZERO; zeta = q.z; }
  Vtx v;
  v.wet = hv > o.h_dry;
  const int b = g.vbc[p];
  if (b >= 0 && sh) hv = dmax(ZERO, sh[b] - q.z);
  v.h = hv;
  v.zeta = zeta;
  return v;
}

// The vertex pass of every stage for vertex p: depths for the wall fluxes and
// wet flags for the cells, from the stage input's water surface (NaN marks a
// dry cell).  first: the pass of a step's first stage, which runs beside that
// step's k_control.  It reads none of what k_control writes: a stopped reach
// is known by `vdone`, and the depths of stage-string vertices (which depend
// on the strings' new stages) are left to k_control.
__device__ __forceinline__ void vertex_stage(const Geo& g, const Opts& o, const double* zeta,
                                             double* hvtx, uint8_t* vwet, const Ctrl* ctrl,
                                             const int* vdone, const double* sh, int first,
                                             bool alldry, int p) {
  if (p >= g.P) return;
  const double2 zb = g.vzb[p];
  const double zv = zb.x;
  const uint64_t bits = (uint64_t)__double_as_longlong(zb.y);
  const int vb = (int)(uint32_t)bits, deg = (int)((bits >> 32) & 0xFF);
  const int r = (int)(bits >> 40);
  if (first ? vdone[r] : ctrl[r].done) return;
  const size_t P = (size_t)g.P;
  double id = ZERO, a = ZERO, zr = zv;
  int counter = 0;
  // in an all-dry chunk every incident cell's zeta is NaN and adds nothing
  const int nread = alldry ? 0 : deg;
  for (int t = 0; t < nread; t += 4) {
    // four cells per round, all loads issued before the sums
    int k[4];
    double w[4], zk[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      k[u] = (t + u < deg) ? g.vell[(t + u) * P + p] : -1;
      w[u] = (t + u < deg) ? g.vwt[(t + u) * P + p] : ZERO;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) zk[u] = zeta[k[u] < 0 ? k[0] : k[u]];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (k[u] < 0 || zk[u] != zk[u]) continue;
      counter++;
      id = id + w[u];
      a = a + w[u] * zk[u];
      zr = dmax(zk[u], zr);
    }
  }
  double zeta_v = (counter > 0) ? a / id : zv;
  zeta_v = dmin(zeta_v, zr);
  double hv = zeta_v - zv;
  if (hv < ZERO) hv = ZERO;
  vwet[p] = hv > o.h_dry;
  if (vb >= 0) {
    if (first) return;
    hv = dmax(ZERO, sh[vb] - zv);
  }
  hvtx[p] = hv;
}

// Every incident cell of chunk c's vertices is dry in the stage input: no
// tile owning one has a wet cell (the flags are written with zeta).
__device__ __forceinline__ bool chunk_dry(const Geo& g, const uint8_t* twet, int c) {
  for (int j = g.vc_ptr[c]; j < g.vc_ptr[c + 1]; ++j)
    if (twet[g.vc_tile[j]]) return false;
  return true;
}

// The vertex pass of stages 2 and up.
__global__ void __launch_bounds__(256)
k_vertex_stage(Geo g, Opts o, const double* zeta, const uint8_t* twet, double* hvtx,
               uint8_t* vwet, const Ctrl* ctrl, const double* sh) {
  const int c = xcd_tile();
  vertex_stage(g, o, zeta, hvtx, vwet, ctrl, nullptr, sh, 0, chunk_dry(g, twet, c),
               c * blockDim.x + threadIdx.x);
}

// The first stage's vertex pass and, as its first blocks (one per reach), the
// step's k_control: the two need only the previous step's last stage and
// write disjoint data, so one launch runs both.
__global__ void __launch_bounds__(256)
k_vertex_first(Geo g, Opts o, const double* zeta, const uint8_t* twet, uint8_t* vwet,
               const int* vdone, CtrlArgs A) {
  const int nr = g.nr;
  if ((int)blockIdx.x < nr) {
    control_block<256>(g, o, A, blockIdx.x);
    return;
  }
  const int c = xcd_map(gridDim.x - nr, blockIdx.x - nr);
  vertex_stage(g, o, zeta, A.hvtx, vwet, A.ctrl, vdone, A.sh, 1, chunk_dry(g, twet, c),
               c * blockDim.x + threadIdx.x);
}

// All vertices after the time loop (solverFVT.f90:483), with zetavtx.  With
// sh null the stage strings' depths are left as interpolated (the final pass
// sets them in k_out_vertex, after the velocity tests that read the
// interpolated depths).
__global__ void __launch_bounds__(256)
k_vertex(Geo g, Opts o, const CellD* d, const uint8_t* dry, double* hvtx, uint8_t* vwet,
         double* zetavtx, const double* sh) {
  const int p = xcd_tile() * blockDim.x + threadIdx.x;
  if (p >= g.P) return;
  const Vtx v = vertex_idw(g, o, d, dry, p, g.vxyz[p], sh);
  vwet[p] = v.wet;
  hvtx[p] = v.h;
  if (zetavtx) zetavtx[p] = v.zeta;
}

// ---------------------------------------------------------------- Green-Gauss pass 1
__global__ void k_gauss1(Geo g, const CellD* d, double* gwork) {
  // gaussgrad.f90:37-42 for zeta, u, v
  const int i = xcd_tile() * blockDim.x + threadIdx.x;
  if (i >= g.N) return;
  const int N = g.N;
  const int n0 = g.gconn[i], n1 = g.gconn[N + i], n2 = g.gconn[2 * N + i];
  const double cx0 = g.gcx[i], cx1 = g.gcx[N + i], cx2 = g.gcx[2 * N + i];
  const double cy0 = g.gcy[i], cy1 = g.gcy[N + i], cy2 = g.gcy[2 * N + i];
  const CellD d0 = ld(&d[n0]), d1 = ld(&d[n1]), d2 = ld(&d[n2]);
  const double z0 = d0.h + g.cs[n0].z, z1 = d1.h + g.cs[n1].z, z2 = d2.h + g.cs[n2].z;
  gwork[i] = cx0 * z0 + cx1 * z1 + cx2 * z2;
  gwork[N + i] = cy0 * z0 + cy1 * z1 + cy2 * z2;
  gwork[2 * N + i] = cx0 * d0.u + cx1 * d1.u + cx2 * d2.u;
  gwork[3 * N + i] = cy0 * d0.u + cy1 * d1.u + cy2 * d2.u;
  gwork[4 * N + i] = cx0 * d0.v + cx1 * d1.v + cx2 * d2.v;
  gwork[5 * N + i] = cy0 * d0.v + cy1 * d1.v + cy2 * d2.v;
}

// limiterBJ.f90:42-81, beta = 1 (cell_grads.f90:59).  The side's ratio only
// enters through min(ratio, 1): where the quotient is at least 1 (its
// numerator is no smaller in size than dq, same sign) that is 1 whatever the
// quotient's bits, so the division is formed only where it can limit.
__device__ __forceinline__ void bj_limit(double q0, const double qn[3], const double rcx[3],
                                         const double rcy[3], double& gx, double& gy) {
  double r[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double qmin = dmin(q0, qn[j]), qmax = dmax(q0, qn[j]);
    const double qj = q0 + gx * rcx[j] + gy * rcy[j];
    const double dq = qj - q0;
    double rr = ONE;
    if (!equals(dq, ZERO)) {
      const double num = dq < 0 ? qmin - q0 : qmax - q0;
      if (dq < 0 ? !(num <= dq) : !(num >= dq)) rr = num / dq;
    }
    const double a = dmin(ONE * rr, ONE), b = dmin(rr, ONE);
    r[j] = dmax(a, b);
  }
  const double ph = dmin(dmin(r[0], r[1]), r[2]);
  gx = ph * gx;
  gy = ph * gy;
}

// ---------------------------------------------------------------- face states
// What the flux loop needs of one cell: its face states at the three sides
// (interior sides as invisc_fluxes.f90:53-81 forms them -- zero for a dry
// cell, depth clipped at 0 -- boundary sides unclipped), the depth gradient
// for the bed-slope term, the modified velocities and the stage input.
struct Faces {
  double fh[3], fu[3], fv[3];
  double dhx, dhy, up, vp, h, umag;  // umag: |u| of the stage input, 0 when dry
  bool dry;
  bool fix;  // hgradfix2 set the gradient: rk hu, hv of slot s-1 are zeroed
  // XSF faces only: fu, fv hold u + du/dx rcx (v likewise) without the
  // y-gradient term, which the xsf edge loop takes from the edge's right
  // cell; rcy of each side and the cell's own du/dy, dv/dy
  double rcy[3], guy, gvy;
};

// find_dry_cells' flags (carried in the state), cell_grads.f90:24-75 (least
// squares of lsq_grad_array.f90:60-73 with the weights and normal-equation
// sums of lsq_setup.f90:105-145 formed on the fly, or Green-Gauss),
// hgradfix2.f90:46-98 and the cell pass of to_vertices.f90:44-55, for cell i.
// vw: the cell's three vertices are wet.
// SKIPDRY (the stage kernel): a dry cell whose sides are interior, wall or
// flux-free sides returns at once with zero faces.  Its interior faces are
// zero whatever its gradients (invisc_fluxes.f90:60-64), the wall routines
// return for dry cells, and its bed-slope term is skipped, so no flux or
// update reads the gradients it would form; the other fields are what the
// full path gives a dry cell.  (The xsf loop reads a dry right cell's
// y-gradients, so it takes the full path.)
template <bool XSF = false, bool SKIPDRY = false>
__device__ __forceinline__ void cell_faces(const Geo& g, const Opts& o, const CellD* din,
                                           const double* gwork, int i, uint64_t meta, int nb0,
                                           int nb1, int nb2, const Tri& tr, bool vw, Faces& F) {
  const int N = g.N;
  const int nb[3] = {nb0, nb1, nb2};
  const CellD di = ld(&din[i]);
  if (SKIPDRY && !XSF && di.dry != ZERO) {
    const uint32_t k0 = slot_bits(meta, 0) & CMS_KIND, k1 = slot_bits(meta, 1) & CMS_KIND,
                   k2 = slot_bits(meta, 2) & CMS_KIND;
    const auto plain = [](uint32_t k) { return k == SK_INTERIOR || k == SK_WALL || k == SK_NONE; };
    if (plain(k0) && plain(k1) && plain(k2)) {
#pragma unroll
      for (int j = 0; j < 3; ++j) { F.fh[j] = ZERO; F.fu[j] = ZERO; F.fv[j] = ZERO; }
      F.dhx = ZERO; F.dhy = ZERO;
      F.up = di.u; F.vp = di.v; F.h = di.h; F.umag = ZERO;
      F.dry = true;
      F.fix = false;
      return;
    }
  }
  const CellS si = ld(&g.cs[i]);
  CellD dn[3];
  CellS sn[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int k = nb[j] < 0 ? i : nb[j];
    dn[j] = ld(&din[k]);
    sn[j] = ld(&g.cs[k]);
  }
  const double hi = di.h, zi = si.z, ui = di.u, vi = di.v;
  const double xci = si.xc, yci = si.yc;
  const bool dry_i = di.dry != ZERO;
  const double zeta_i = hi + zi;
  double nzeta[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) nzeta[j] = dn[j].h + sn[j].z;
  double gzx = ZERO, gzy = ZERO, gux = ZERO, guy = ZERO, gvx = ZERO, gvy = ZERO;
  if (o.limiter != 0) {
    if (o.gradient == STORM_GRAD_GAUSS) {
      // gaussgrad.f90:45-52
      const double* w = gwork;
      const int n0 = g.gconn[i], n1 = g.gconn[N + i], n2 = g.gconn[2 * N + i];
      const double w0 = g.gwt[i], w1 = g.gwt[N + i], w2 = g.gwt[2 * N + i];
      gzx = HALF * (w[i] + w0 * w[n0] + w1 * w[n1] + w2 * w[n2]);
      gzy = HALF * (w[N + i] + w0 * w[N + n0] + w1 * w[N + n1] + w2 * w[N + n2]);
      gux = HALF * (w[2 * N + i] + w0 * w[2 * N + n0] + w1 * w[2 * N + n1] + w2 * w[2 * N + n2]);
      guy = HALF * (w[3 * N + i] + w0 * w[3 * N + n0] + w1 * w[3 * N + n1] + w2 * w[3 * N + n2]);
      gvx = HALF * (w[4 * N + i] + w0 * w[4 * N + n0] + w1 * w[4 * N + n1] + w2 * w[4 * N + n2]);
      gvy = HALF * (w[5 * N + i] + w0 * w[5 * N + n0] + w1 * w[5 * N + n1] + w2 * w[5 * N + n2]);
    } else {
      // ACTIVATE HDLS: the t2tHD molecule for every cell, unit weights
      // (hdlsq_setup.f90:39-74 with type 1, hdlsq_grad_array.f90:36-48)
      const bool hd = o.gradient == STORM_GRAD_HDLS;
      double dz = ZERO, ez = ZERO, du = ZERO, eu = ZERO, dv = ZERO, ev = ZERO;
      double a = ZERO, b = ZERO, cc = ZERO;
      auto member = [&](double mxc, double myc, double mzeta, double mu, double mv) {
        const double dx = mxc - xci, dy = myc - yci;
        const double w = ONE / (hd ? ONE : dx * dx + dy * dy);
        const double fz = mzeta - zeta_i;
        const double fu = mu - ui;
        const double fv = mv - vi;
        a = a + w * dx * dx; b = b + w * dx * dy; cc = cc + w * dy * dy;
        dz = dz + w * fz * dx; ez = ez + w * fz * dy;
        du = du + w * fu * dx; eu = eu + w * fu * dy;
        dv = dv + w * fv * dx; ev = ev + w * fv * dy;
      };
      const bool interior = !hd && ((slot_bits(meta, 0) | slot_bits(meta, 1) |
                                     slot_bits(meta, 2)) & CMS_KIND) == SK_INTERIOR;
      if (interior) {
        // elemnt_db.f90:50-68: a cell with three edge neighbours keeps them,
        // in edge-number order, which is the processing order of its sides
#pragma unroll
        for (int tt = 0; tt < 3; ++tt) {
          const int j = (int)(meta >> (2 * tt)) & 3;
          const double mxc = (j == 0) ? sn[0].xc : (j == 1) ? sn[1].xc : sn[2].xc;
          const double myc = (j == 0) ? sn[0].yc : (j == 1) ? sn[1].yc : sn[2].yc;
          const double mz = (j == 0) ? nzeta[0] : (j == 1) ? nzeta[1] : nzeta[2];
          const double mu = (j == 0) ? dn[0].u : (j == 1) ? dn[1].u : dn[2].u;
          const double mv = (j == 0) ? dn[0].v : (j == 1) ? dn[1].v : dn[2].v;
          member(mxc, myc, mz, mu, mv);
        }
      } else {
        const int t0 = g.mptr[i], t1 = g.mptr[i + 1];
        for (int t = t0; t < t1; ++t) {
          const int k = g.mcell[t];
          const CellD dk = ld(&din[k]);
          const CellS sk = ld(&g.cs[k]);
          member(sk.xc, sk.yc, dk.h + sk.z, dk.u, dk.v);
        }
      }
      const double d = a * cc - b * b;
      const Recip rd = recip(d, cc * dz - b * ez);
      gzx = qdiv(rd, cc * dz - b * ez, d); gzy = qdiv(rd, a * ez - dz * b, d);
      gux = qdiv(rd, cc * du - b * eu, d); guy = qdiv(rd, a * eu - du * b, d);
      gvx = qdiv(rd, cc * dv - b * ev, d); gvy = qdiv(rd, a * ev - dv * b, d);
    }
  }
  double rcx[3], rcy[3];
  {
    const Side s0 = side<0>(tr), s1 = side<1>(tr), s2 = side<2>(tr);
    rcx[0] = s0.mx - xci; rcx[1] = s1.mx - xci; rcx[2] = s2.mx - xci;
    rcy[0] = s0.my - yci; rcy[1] = s1.my - yci; rcy[2] = s2.my - yci;
  }
  if (o.limiter == 2) {
    double qz[3], qu[3], qv[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      qz[j] = (nb[j] < 0) ? zeta_i : nzeta[j];
      qu[j] = (nb[j] < 0) ? ui : dn[j].u;
      qv[j] = (nb[j] < 0) ? vi : dn[j].v;
    }
    bj_limit(zeta_i, qz, rcx, rcy, gzx, gzy);
    bj_limit(ui, qu, rcx, rcy, gux, guy);
    bj_limit(vi, qv, rcx, rcy, gvx, gvy);
  }
  // bed_slope_VTX.f90:32-33
  const double area = tri_area(tr);
  const double nzx = HALF * (tr.zv[0] * (tr.y[1] - tr.y[2]) + tr.zv[1] * (tr.y[2] - tr.y[0]) +
                             tr.zv[2] * (tr.y[0] - tr.y[1]));
  const double nzy = HALF * (tr.zv[0] * (tr.x[2] - tr.x[1]) + tr.zv[1] * (tr.x[0] - tr.x[2]) +
                             tr.zv[2] * (tr.x[1] - tr.x[0]));
  const Recip ra = recip(area, nzx);
  const double dzx = qdiv(ra, nzx, area), dzy = qdiv(ra, nzy, area);
  double dhx = gzx - dzx, dhy = gzy - dzy;
  double up = ui, vp = vi;
  bool fix = false;
  if (!dry_i) {
    // hgradfix2.f90:49-96 (zvtx of the highest vertex: dataprepFVT.f90:293-306)
    const double zhigh = dmax(dmax(tr.zv[0], tr.zv[1]), tr.zv[2]);
    bool setgrad = zeta_i < zhigh;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (nb[k] < 0) continue;
      if (dn[k].dry != ZERO && sn[k].z < zeta_i) setgrad = false;
    }
    if (setgrad) {
      dhx = -dzx; dhy = -dzy;
      up = ZERO; vp = ZERO;
      fix = true;
      gux = ZERO; guy = ZERO; gvx = ZERO; gvy = ZERO;
    }
    // to_vertices.f90:44-55
    if (!vw) { up = ZERO; vp = ZERO; }
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double fh = hi + dhx * rcx[j] + dhy * rcy[j];
    double fu, fv;
    if (XSF) {  // the partial sums; the edge adds a y-gradient term
      fu = up + gux * rcx[j];
      fv = vp + gvx * rcx[j];
    } else {
      fu = up + gux * rcx[j] + guy * rcy[j];
      fv = vp + gvx * rcx[j] + gvy * rcy[j];
    }
    if (nb[j] >= 0) {
      // a dry cell: velocity 0 and depth 0, or under invisc_fluxes_xsf.f90:
      // 56-58 its cell depth
      if (dry_i) { fh = XSF ? hi : ZERO; fu = ZERO; fv = ZERO; }
      else if (fh < ZERO) fh = ZERO;
    }
    F.fh[j] = fh; F.fu[j] = fu; F.fv[j] = fv;
    if (XSF) F.rcy[j] = rcy[j];
  }
  if (XSF) { F.guy = guy; F.gvy = gvy; }
  F.dhx = dhx; F.dhy = dhy; F.up = up; F.vp = vp; F.h = hi;
  F.umag = dry_i ? ZERO : sqrt(ui * ui + vi * vi);
  F.dry = dry_i;
  F.fix = fix;
}

// The local dt of the reference's last cell from the stage input, which the
// steady-mode implicit friction uses (solverFVT.f90:260-265,336).
__device__ double last_cell_dt(const Geo& g, const Opts& o, const CellD* din, int i) {
  const CellD d = ld(&din[i]);
  const Tri tr = load_tri(g, i);
  double umg = ZERO;
  const bool dr = find_dry(d.h, d.u, d.v, o.h_dry, o.h_wet, &umg);
  const double wc = dr ? ZERO : sqrt(G * d.h);
  double dd = umg + wc;
  if (dd < SMALL) dd = 1.044;
  const double perim = side<0>(tr).len + side<1>(tr).len + side<2>(tr).len;
  return o.cfl * tri_area(tr) / (dd * perim);
}

// Flux through one boundary side of cell i (the routines of
// invisc_fluxes.f90:195-238), as the change of phi(i,:).  h1, h2: vertex
// depths at the ends of the edge in the reference's edge direction; sq, sh:
// the current values of every string; q_*: the inflowQbyH results of the
// reach's strings (index local to the reach).
__device__ void boundary_flux(const Opts& o, const RInfo& ri, const double* sq, const double* sh,
                              const double* q_at, const int* q_null, const int* q_skip,
                              uint32_t bits, const Side& sd, double fh, double fu, double fv,
                              double hi, bool dry_i, double h1, double h2, double& d0, double& d1,
                              double& d2) {
  const uint32_t kind = bits & CMS_KIND;
  const double h_dry = o.h_dry;
  const double nx = sd.nx, ny = sd.ny, L = sd.len;
  d0 = ZERO; d1 = ZERO; d2 = ZERO;
  if (kind == SK_WALL) {
    if (dry_i) return;
    if (o.btype == 0) {
      // wenslp3_fluxes.f90:40-61
      double he = sqrt((h1 * h1 + h1 * h2 + h2 * h2) * ONE_THIRD);
      if (he < ZERO) he = ZERO;
      if (he < h_dry) return;
      d1 = -(HALF * G * he * he * nx * L);
      d2 = -(HALF * G * he * he * ny * L);
    } else {
      // weslp2_fluxes.f90:40-75
      if (fh < h_dry) return;
      double urr = fu * sd.tx + fv * sd.ty;
      urr = urr + 2.0 * sqrt(G * fh);
      const double hrr = urr * urr / 4.0 / G;
      const double f1 = HALF * G * hrr * hrr * nx, f2 = HALF * G * hrr * hrr * ny;
      const double dh = ZERO;
      d1 = -((f1 + dh * nx) * L);
      d2 = -((f2 + dh * ny) * L);
    }
    return;
  }
  const int ch = (bits >> CMS_CHAIN) & 0xFF;
  if (kind == SK_IN_Q1 || kind == SK_IN_Q5) {
    // inflowQbyH.f90:131-210 (vtype 1) / :321-400 (vtype 5)
    if (q_skip[ch]) return;
    if (kind == SK_IN_Q5 && dry_i) return;
    const double qin = sq[ri.in_base + ch];
    const double atotal = q_at[ch];
    double hl = fh, ue, ve;
    if (q_null[ch]) hl = ri.bch_max;
    if (hl < h_dry) { hl = dmax(hl, ZERO); ue = ZERO; ve = ZERO; }
    else { ue = fu; ve = fv; }
    const double qfrac = qin / atotal;
    double urr = -qfrac;
    const double ul = ue * nx + ve * ny;
    const double sq = sqrt(G * hl);
    double hr = 2.0 * sq + ul - urr;
    hr = hr * hr / 4.0 / G;
    urr = -qfrac * nx;
    const double vrr = -qfrac * ny;
    if (hr < h_dry) return;
    const double f0 = hr * (urr * nx + vrr * ny);
    const double f1 = hr * urr * (urr * nx + vrr * ny) + HALF * G * hr * hr * nx;
    const double f2 = hr * vrr * (urr * nx + vrr * ny) + HALF * G * hr * hr * ny;
    d0 = -(f0 * L); d1 = -(f1 * L); d2 = -(f2 * L);
  } else if (kind == SK_IN_H) {
    // inflowH.f90:31-91
    double hl = dmax(fh, ZERO), ue, ve;
    if (hi < h_dry) { ue = ZERO; ve = ZERO; } else { ue = fu; ve = fv; }
    const double hr = dmax(ZERO, sq[ri.in_base + ch] - sd.zm);
    const double ul = ue * nx + ve * ny;
    const double sl = sqrt(G * hl), sr = sqrt(G * hr);
    double urr = ul + (sl - sr);
    const double vrr = urr * ny;
    urr = urr * nx;
    if (hl < h_dry && hr < h_dry) return;
    double f[3];
    flux_ac(nx, ny, hl, ue, ve, hr, urr, vrr, f);
    d0 = -(f[0] * L); d1 = -(f[1] * L); d2 = -(f[2] * L);
  } else if (kind == SK_OUT_H) {
    // outflowbyH.f90:33-135
    double hl = fh, ue, ve;
    if (dry_i || hl < h_dry) { ue = ZERO; ve = ZERO; hl = dmax(hl, ZERO); }
    else { ue = fu; ve = fv; }
    double hr = dmax(ZERO, sh[ri.out_base + ch] - sd.zm);
    if (hl < h_dry && hr < h_dry) return;
    double ul, vl, urr, vrr;
    if (hr < h_dry) { urr = ZERO; ul = ue; vrr = ZERO; vl = ve; }
    else {
      ul = ue * nx + ve * ny;
      vl = ue * sd.tx + ve * sd.ty;
      const double sl = sqrt(G * hl), sr = sqrt(G * hr);
      urr = ul + (sl - sr);
      vrr = vl;
      ul = ue;
      vl = ve;
      const double a = urr, b = vrr;
      urr = a * nx + b * sd.tx;
      vrr = a * ny + b * sd.ty;
    }
    bool set_wall = false;
    if (o.hvalve) {
      const double dot = urr * nx + vrr * ny;
      if (dot < ZERO) { set_wall = true; hr = hl; urr = ZERO; vrr = ZERO; }
    }
    double f[3];
    flux_ac(nx, ny, hl, ul, vl, hr, urr, vrr, f);
    if (set_wall) f[0] = ZERO;
    d0 = -(f[0] * L); d1 = -(f[1] * L); d2 = -(f[2] * L);
  } else if (kind == SK_OUT_CRIT) {
    // crit_outflow.f90:32-105
    double hl = fh, ue, ve;
    if (dry_i || hl < h_dry) { ue = ZERO; ve = ZERO; hl = dmax(hl, ZERO); }
    else { ue = fu; ve = fv; }
    if (hl < h_dry) return;
    const double ul = ue * nx + ve * ny, vl = ue * sd.tx + ve * sd.ty;
    double urr = ul, vrr = vl;
    double hr = urr * urr / G;
    const double a = urr, b = vrr;
    urr = a * nx + b * sd.tx;
    vrr = a * ny + b * sd.ty;
    if (o.hvalve) {
      const double dot = urr * nx + vrr * ny;
      if (dot < ZERO) { hr = hl; urr = ZERO; vrr = ZERO; }
    }
    const double f0 = hr * (urr * nx + vrr * ny);
    const double f1 = hr * urr * (urr * nx + vrr * ny) + HALF * G * hr * hr * nx;
    const double f2 = hr * vrr * (urr * nx + vrr * ny) + HALF * G * hr * hr * ny;
    const double dh = ZERO;
    d0 = -(f0 * L); d1 = -((f1 + dh * nx) * L); d2 = -((f2 + dh * ny) * L);
  } else if (kind == SK_OUT_FREE) {
    // free_flow.f90:33-90
    const double hl = dmax(fh, ZERO);
    if (hl < h_dry) return;
    double ue, ve;
    if (hi < h_dry) { ue = ZERO; ve = ZERO; } else { ue = fu; ve = fv; }
    double urr = ue, vrr = ve;
    bool set_wall = false;
    if (o.hvalve) {
      const double dot = urr * nx + vrr * ny;
      if (dot < 0) { set_wall = true; urr = ZERO; vrr = ZERO; }
    }
    double f[3];
    flux_ac(nx, ny, hl, ue, ve, hl, urr, vrr, f);
    if (set_wall) f[0] = ZERO;
    d0 = -(f[0] * L); d1 = -(f[1] * L); d2 = -(f[2] * L);
  }
}

// Neighbour data an interior side needs: its face state at the shared side,
// its bed, stage-input depth and dry flag.
struct Nbr {
  double fh, fu, fv, z, h;
  bool dry;
};

// invisc_fluxes.f90:39-166 for one interior side of cell i: the flux of the
// edge computed exactly as the reference does for it (left state = its e(1)
// cell, normal out of e(1)), returned as cell i's change of phi.  Returns
// false where the edge carries no flux (zero change, +0 for both cells);
// otherwise the neighbour's change is the exact negation of cell i's.
__device__ __forceinline__ bool interior_flux(const Opts& o, uint32_t bits, const Side& sd,
                                              bool dry_i, double hi, double zi, double fho,
                                              double fuo, double fvo, const Nbr& nb, double& d0,
                                              double& d1, double& d2) {
  d0 = ZERO; d1 = ZERO; d2 = ZERO;
  if (dry_i && nb.dry) return false;
  const double h_dry = o.h_dry, h_wet = o.h_wet;
  const bool left = (bits >> CMS_LEFT) & 1;
  double hl = left ? fho : nb.fh, ul = left ? fuo : nb.fu, vl = left ? fvo : nb.fv;
  double hr = left ? nb.fh : fho, ur = left ? nb.fu : fuo, vr = left ? nb.fv : fvo;
  const bool dL = left ? dry_i : nb.dry, dR = left ? nb.dry : dry_i;
  const double zL = left ? zi : nb.z, zR = left ? nb.z : zi;
  const double wL = left ? hi : nb.h, wR = left ? nb.h : hi;
  const double nx = left ? sd.nx : -sd.nx, ny = left ? sd.ny : -sd.ny;
  if (hl < h_dry && hr < h_dry) return false;
  bool set_wall = false;
  if (dL || dR) {
    const double h0 = dmax(hr, hl) + sd.zm;
    if (dL && zL + h_wet > h0) { set_wall = true; if (!o.xsf) hl = hr; }
    if (dR && zR + h_wet > h0) { set_wall = true; if (!o.xsf) hr = hl; }
  }
  double f[3];
  if (o.xsf) {
    // invisc_fluxes_xsf.f90:118-121: flux_ac of the states, mass flux
    // zeroed at an ALGORITHM I1 wall
    flux_ac(nx, ny, hl, ul, vl, hr, ur, vr, f);
    if (set_wall) f[0] = ZERO;
  } else if (set_wall) {
    f[0] = ZERO;
    hl = dmax(hl, hr);
    f[1] = HALF * G * hl * hl * nx;
    f[2] = HALF * G * hl * hl * ny;
  } else if (o.flux == STORM_FLUX_AC) {
    flux_ac(nx, ny, hl, ul, vl, hr, ur, vr, f);
  } else if (o.flux == STORM_FLUX_RUSANOV) {
    const double cl = dL ? ZERO : sqrt(G * wL), cr = dR ? ZERO : sqrt(G * wR);
    flux_rusanov(nx, ny, hl, ul, vl, cl, hr, ur, vr, cr, f);
  } else if (o.flux == STORM_FLUX_HLL) {
    flux_hll(nx, ny, hl, ul, vl, hr, ur, vr, h_dry, f);
  } else {
    flux_bgnvc(nx, ny, hl, ul, vl, hr, ur, vr, f);
  }
  const double L = sd.len;
  if (left) { d0 = -(f[0] * L); d1 = -(f[1] * L); d2 = -(f[2] * L); }
  else { d0 = f[0] * L; d1 = f[1] * L; d2 = f[2] * L; }
  return true;
}

// ---------------------------------------------------------------- stage kernel
// Shared-memory layout of one tile (CAP cells + halo).  The inflowQbyH
// buffers of a string (QMAX edges) overlay the cell arrays, which the fixed
// point runs before.
constexpr int QMAX = 512;  // edges of one inflowQbyH string

struct Lds {
  double *fh, *fu, *fv;          // [3][CAP] face states, side-major
  double *rcy, *gy;              // XSF only: [3][CAP] rcy per side, [2][CAP] du/dy, dv/dy
  double *z, *h;                 // [CAP] bed and stage-input depth
  double *vx, *vy, *vz;          // [3][TMAX] vertex coordinates and bed of the owned cells
  uint8_t* dry;                  // [CAP] dry flags
  double *qh, *qu, *qv, *qa, *qq;  // [QMAX] string edges (overlay)
  uint8_t* qok;                    // [QMAX] (overlay)
};

// Face states, bed and depth for every cell of the tile and halo; vertex
// coordinates and bed for the owned cells only: 38.0 KB at capacity 256
// (four blocks per CU, as many as the registers allow).
__host__ __device__ constexpr size_t lds_bytes(int cap, bool xsf = false) {
  return sizeof(double) * ((xsf ? 16 : 11) * (size_t)cap + 9 * (size_t)tile_tmax(cap)) +
         (size_t)cap;
}

__device__ __forceinline__ Lds lds_layout(double* base, int cap, bool xsf = false) {
  Lds L;
  L.fh = base; L.fu = L.fh + 3 * cap; L.fv = L.fu + 3 * cap;
  L.z = L.fv + 3 * cap; L.h = L.z + cap;
  const int tm = tile_tmax(cap);
  L.vx = L.h + cap; L.vy = L.vx + 3 * tm; L.vz = L.vy + 3 * tm;
  // XSF kernels only (lds_bytes(cap, true)): rcy and the y-gradients
  L.rcy = L.vz + 3 * tm;
  L.gy = L.rcy + 3 * cap;
  L.dry = reinterpret_cast<uint8_t*>(xsf ? L.gy + 2 * cap : L.vz + 3 * tm);
  // overlay of the 20*cap cell doubles: 5*QMAX doubles and QMAX flags
  L.qh = base; L.qu = L.qh + QMAX; L.qv = L.qu + QMAX; L.qa = L.qv + QMAX; L.qq = L.qa + QMAX;
  L.qok = reinterpret_cast<uint8_t*>(L.qq + QMAX);
  return L;
}
static_assert(5 * QMAX * 8 + QMAX <= 11 * 256 * 8, "string buffers must fit the cell arrays");

__device__ __forceinline__ int lid(uint64_t w, int j) { return (int)(w >> (16 * j)) & 0xFFFF; }

// side j (known only at run time) of tile cell t
template <int CAP>
__device__ __forceinline__ Side side_dyn(const Lds& L, int t, int j) {
  constexpr int TM = tile_tmax(CAP);
  const int a = j * TM + t, b = (j == 2 ? 0 : j + 1) * TM + t;
  return side_of(L.vx[a], L.vy[a], L.vz[a], L.vx[b], L.vy[b], L.vz[b]);
}

// inflowQbyH.f90:31-126 (vtype 1) / :221-316 (vtype 5): the fixed point on
// the wetted length of string ch, by the whole block (edge l of the string
// is held by thread l % CAP).  The face state of each string edge's cell is
// formed here from the stage input, as the cell's own tile forms it.  Every
// sum runs over the edges in string order (the reference's sequential sum);
// all threads form the sums from shared memory, so the loop exit is uniform
// across the block.
template <int CAP>
__device__ void bcq_string(const Geo& g, const Opts& o, const CellD* din, const double* gwork,
                           const uint8_t* vwet, const double* sq, const RInfo& ri, int ch,
                           const Lds& L, double* s_at, int* s_null, int* s_skip) {
  const int t = threadIdx.x;
  const int gch = ri.in_base + ch;  // the string's index among all strings
  const int vt = g.in_type[gch];
  const int e0 = g.qs_ptr[gch], n = g.qs_ptr[gch + 1] - e0;
  const double qin = sq[gch];
  if (vt != 1 && vt != 5) return;
  if (qin < DBL_MIN) {
    if (t == 0) s_skip[ch] = 1;
    __syncthreads();
    return;
  }
  const double h_dry = o.h_dry;
  // qok: bit 1 the edge takes part (vtype 5 skips dry cells), bit 0 its
  // terms enter the current sums
  for (int l = t; l < n; l += CAP) {
    const int e = e0 + l;
    const int i = g.qs_cell[e], m = g.qs_slot[e];
    const Tri tr = load_tri(g, i);
    const int N = g.N;
    const bool vw = vwet[g.cvtx[i]] && vwet[g.cvtx[N + i]] && vwet[g.cvtx[2 * N + i]];
    Faces F;
    cell_faces(g, o, din, gwork, i, g.cmeta[i], g.cnbr[i], g.cnbr[N + i], g.cnbr[2 * N + i], tr,
               vw, F);
    const double fh = (m == 0) ? F.fh[0] : (m == 1) ? F.fh[1] : F.fh[2];
    L.qh[l] = fh;
    L.qu[l] = (m == 0) ? F.fu[0] : (m == 1) ? F.fu[1] : F.fu[2];
    L.qv[l] = (m == 0) ? F.fv[0] : (m == 1) ? F.fv[1] : F.fv[2];
    const bool act = !(vt == 5 && F.dry);
    // initial wetted length (inflowQbyH.f90:40-56)
    const double hl = F.dry ? ZERO : dmax(fh, ZERO);
    L.qa[l] = hl * g.qs_len[e];
    L.qok[l] = act ? 3 : 0;  // bit 1: takes part in every pass
  }
  __syncthreads();
  double atotal = ZERO;
  for (int k = 0; k < n; ++k)
    if (L.qok[k] & 1) atotal = atotal + L.qa[k];
  const bool null_area = equals(atotal, ZERO);
  int counter = 0;
  double newatotal, qtotal;
  for (;;) {
    __syncthreads();
    const double qfrac = qin / atotal;
    for (int l = t; l < n; l += CAP) {
      const int e = e0 + l;
      const int act = L.qok[l] & 2;
      int ok = 0;
      if (act) {
        double hl = L.qh[l], ue, ve;
        if (null_area) hl = ri.bch_max;
        if (hl < h_dry) { hl = dmax(hl, ZERO); ue = ZERO; ve = ZERO; }
        else { ue = L.qu[l]; ve = L.qv[l]; }
        const double nx = g.qs_nx[e], ny = g.qs_ny[e], len = g.qs_len[e];
        double ur = -qfrac;
        const double ul = ue * nx + ve * ny;
        const double sq = sqrt(G * hl);
        double hr = 2.0 * sq + ul - ur;
        hr = hr * hr / 4.0 / G;
        ur = -qfrac * nx;
        const double vr = -qfrac * ny;
        if (!(hr < h_dry)) {
          L.qa[l] = hr * len;
          L.qq[l] = (ur * nx + vr * ny) * hr * len;
          ok = 1;
        }
      }
      L.qok[l] = (uint8_t)(act | ok);
    }
    __syncthreads();
    newatotal = ZERO;
    qtotal = ZERO;
    for (int k = 0; k < n; ++k)
      if (L.qok[k] & 1) { newatotal = newatotal + L.qa[k]; qtotal = qtotal + L.qq[k]; }
    qtotal = fabs(qtotal);
    if (fabs(qtotal - qin) / qin < 0.005 || counter > 50) break;
    counter++;
    atotal = newatotal;
  }
  if (t == 0) {
    s_at[ch] = newatotal;
    s_null[ch] = null_area;
    s_skip[ch] = 0;
  }
  __syncthreads();
}

// One RK stage of one tile (solverFVT.f90:158-331), after k_vertex has
// formed the vertex depths and wet flags of the stage:
//   2. the inflowQbyH fixed points of the strings the tile touches;
//   3. the face states of the tile and halo cells (find_dry_cells' flags,
//      cell_grads, hgradfix2, to_vertices' cell pass) into shared memory;
//   4. per tile cell: every flux of its three sides, gathered in the
//      reference's accumulation order (no atomics), bed slope, explicit
//      friction if chosen, SSP-RK update, clip and wet/dry recovery; in the
//      last stage also implicit friction, hvalve, the residual partial sums
//      and the next step's CFL candidate (per-tile partials, reduced in a
//      fixed order by k_control).
// Four waves per SIMD at capacity 256 (at most 128 VGPRs, under 40 KB of
// shared memory): the inflow strings run before the other loads are issued
// and each cell's rk slots and record are loaded in phase 4, so no value is
// held in registers through the strings or phases 2-3.  The few registers the
// strings' fixed point spills are reloaded on that path alone.
// MODE 0 runs it whole.  Residual smoothing splits it around k_rsavg:
// MODE 1 stops after the residual (fluxes and sources) of every tile cell
// and writes it to S.phi0; MODE 2 skips phases 2-3 and the residual, reads
// the smoothed residual from S.phis and runs the rest of phase 4.
template <int CAP, bool XSF = false, int MODE = 0>
__global__ void __launch_bounds__(CAP) __attribute__((amdgpu_waves_per_eu(STAGE_WAVES(CAP, XSF))))
k_stage(Geo g, Opts o, Tiles tl, Stage st, State S) {
  // fluxes once per edge from the side records (the xsf loop forms them per
  // cell side, as before)
  constexpr bool EDGES = !XSF && MODE != 2;
  extern __shared__ double lds_dyn[];
  __shared__ double s_at[MAX_STRINGS];
  __shared__ int s_null[MAX_STRINGS], s_skip[MAX_STRINGS];
  __shared__ double s_dlast;
  const int4 hd0 = tl.hdr[2 * blockIdx.x], hd1 = tl.hdr[2 * blockIdx.x + 1];
  const int c0 = hd0.x, T = hd0.y, hb = hd0.z, H = hd0.w, b = hd1.z;
  const Ctrl* c = &S.ctrl[hd1.w];
  const int t = threadIdx.x;
  const int N = g.N;

  const uint64_t qmask = (uint64_t)(uint32_t)hd1.x | ((uint64_t)(uint32_t)hd1.y << 32);
  const Lds L = lds_layout(lds_dyn, CAP, XSF);

  // 2. inflow discharge strings (their buffers overlay the cell arrays), in
  // the few tiles that touch one, before any other load is issued
  if (MODE != 2 && qmask) {
    const RInfo rq = g.reach[hd1.w];
    const int dq = c->done;
    if (st.s == 1 && t == 0 && b == rq.tile0) S.vdone[hd1.w] = dq;
    if (dq) return;
    for (uint64_t qm = qmask; qm; qm &= qm - 1)
      bcq_string<CAP>(g, o, S.din, S.gwork, S.vwet, S.sq, rq, __builtin_ctzll(qm), L, s_at,
                      s_null, s_skip);
  }

  // Loads that need only the tile header are issued next, so their latency
  // overlaps the wait for the reach's control block: an owned cell's
  // metadata and local neighbour ids and, for every tile and halo cell, the
  // neighbour and vertex ids phase 3 gathers through.
  uint64_t pf_meta = 0, pf_ln = 0;
  int ci = 0, nb0 = -1, nb1 = -1, nb2 = -1, v0 = 0, v1 = 0, v2 = 0;
  if (t < T) {
    const int i = c0 + t;
    pf_meta = g.cmeta[i];
    pf_ln = tl.cln[i];
    ci = i;
    nb0 = g.cnbr[i]; nb1 = g.cnbr[N + i]; nb2 = g.cnbr[2 * N + i];
    v0 = g.cvtx[i]; v1 = g.cvtx[N + i]; v2 = g.cvtx[2 * N + i];
  } else if (MODE != 2 && t < T + H) {
    const int4 r0 = tl.hrec[2 * (hb + t - T)], r1 = tl.hrec[2 * (hb + t - T) + 1];
    ci = r0.x; nb0 = r0.y; nb1 = r0.z; nb2 = r0.w;
    v0 = r1.x; v1 = r1.y; v2 = r1.z;
    pf_meta = tl.hmeta[hb + t - T];
  }
  const RInfo ri = g.reach[hd1.w];
  const int done = c->done;
  if (st.s == 1 && t == 0 && b == ri.tile0) S.vdone[hd1.w] = done;
  if (done) return;
  const double h_dry = o.h_dry, h_wet = o.h_wet;
  if (MODE != 1 && t == 0 && o.steady && st.last) {
    s_dlast = last_cell_dt(g, o, S.din, ri.last_cell);
    // k_control advances the reach's ttime by it
    if (b == ri.tile0) S.ctrl[hd1.w].delt_last = s_dlast;
  }

  Faces F;
  int tile_wet = 1;  // some tile or halo cell is wet (phase 3b)
  if constexpr (MODE == 2) {
    // what phase 4 reads of the faces: the stage input's depth, dry flag
    // and |u| (hgradfix2's zeroed slot values come from the slots themselves)
    if (t < T) {
      const CellD di = ld(&S.din[ci]);
      F.h = di.h;
      F.dry = di.dry != ZERO;
      F.umag = F.dry ? ZERO : sqrt(di.u * di.u + di.v * di.v);
    }
  } else {
  // 3. face states of the tile and halo cells
  if (t < T + H) {
    const bool own = t < T;
    const int i = ci;
    const uint64_t meta = pf_meta;
    Tri tr;
    {
      const double4 q0 = g.vxyz[v0], q1 = g.vxyz[v1], q2 = g.vxyz[v2];
      tr.x[0] = q0.x; tr.y[0] = q0.y; tr.zv[0] = q0.z;
      tr.x[1] = q1.x; tr.y[1] = q1.y; tr.zv[1] = q1.z;
      tr.x[2] = q2.x; tr.y[2] = q2.y; tr.zv[2] = q2.z;
    }
    const bool vw = S.vwet[v0] && S.vwet[v1] && S.vwet[v2];
    cell_faces<XSF, true>(g, o, S.din, S.gwork, i, meta, nb0, nb1, nb2, tr, vw, F);
    if (own) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        constexpr int TM = tile_tmax(CAP);
        L.vx[k * TM + t] = tr.x[k]; L.vy[k * TM + t] = tr.y[k]; L.vz[k * TM + t] = tr.zv[k];
      }
    }
    if (own && F.fix) {
      st.hin[N + i] = ZERO;  // rku, rkv(i, rkstep-1) (hgradfix2.f90:80-81)
      st.hin[2 * N + i] = ZERO;
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      L.fh[j * CAP + t] = F.fh[j]; L.fu[j * CAP + t] = F.fu[j]; L.fv[j * CAP + t] = F.fv[j];
      if (XSF) L.rcy[j * CAP + t] = F.rcy[j];
    }
    if (XSF) { L.gy[t] = F.guy; L.gy[CAP + t] = F.gvy; }
    L.z[t] = g.cs[i].z;
    L.h[t] = F.h;
    L.dry[t] = F.dry;
  }
  if constexpr (EDGES) {
    // 3b. the flux of every side of the tile's owned cells, once per edge,
    // one record per thread.  The two face states of an edge are read only
    // by its record, which overwrites them with the two cells' changes of
    // phi (the neighbour's is the exact negation, +0 where the edge carries
    // no flux), so phase 4 reads its sides' changes where its faces were.
    // A tile without a wet cell has no flux across its interior sides and
    // walls: only its other boundary sides are formed, and phase 4 takes
    // zero for the rest.
    tile_wet = __syncthreads_or(t < T + H && !F.dry);
    const int2 eh = tl.ehdr[blockIdx.x];
    const int n_int = eh.y & 0xFFFF, n_all = n_int + (eh.y >> 16);
    if (tile_wet || n_all > n_int) {
    for (int e = (tile_wet ? 0 : n_int) + t; e < n_all; e += CAP) {
      const uint64_t r = tl.erec[eh.x + e];
      const int a = (int)(r & 0xFFFF), ja = (int)(r >> 16) & 3;
      const uint32_t bits = (uint32_t)(r >> 18) & ((1u << CM_SLOT_BITS) - 1);
      const uint32_t kind = bits & CMS_KIND;
      const int sa = ja * CAP + a;
      const bool dry_a = L.dry[a];
      double d0 = ZERO, d1 = ZERO, d2 = ZERO;
      if (kind == SK_INTERIOR) {
        const int k = (int)(r >> 36) & 0xFFFF, sk = ((bits >> CMS_NL) & 3) * CAP + k;
        Nbr nb;
        nb.dry = L.dry[k];
        bool formed = false;
        if (!(dry_a && nb.dry)) {
          nb.fh = L.fh[sk]; nb.fu = L.fu[sk]; nb.fv = L.fv[sk]; nb.z = L.z[k]; nb.h = L.h[k];
          formed = interior_flux(o, bits, side_dyn<CAP>(L, a, ja), dry_a, L.h[a], L.z[a], L.fh[sa],
                                 L.fu[sa], L.fv[sa], nb, d0, d1, d2);
        }
        if (k < T) {
          L.fh[sk] = formed ? -d0 : ZERO; L.fu[sk] = formed ? -d1 : ZERO; L.fv[sk] = formed ? -d2 : ZERO;
        }
      } else if (!(dry_a && kind == SK_WALL)) {
        const int i = c0 + a;
        const int va = g.cvtx[ja * N + i], vb = g.cvtx[(ja == 2 ? 0 : ja + 1) * N + i];
        const double ha = S.hvtx[va], hb2 = S.hvtx[vb];
        const bool rev = (bits >> CMS_REV) & 1;
        boundary_flux(o, ri, S.sq, S.sh, s_at, s_null, s_skip, bits, side_dyn<CAP>(L, a, ja),
                      L.fh[sa], L.fu[sa], L.fv[sa], L.h[a], dry_a, rev ? hb2 : ha, rev ? ha : hb2,
                      d0, d1, d2);
      }
      L.fh[sa] = d0; L.fu[sa] = d1; L.fv[sa] = d2;
    }
    __syncthreads();
    }
  } else {
    __syncthreads();
  }
  }  // MODE != 2

  // 4. update of the tile cells
  double top = ZERO, bot = ZERO, mxphi = ZERO, cand = DBL_MAX, topq = ZERO, botq = ZERO;
  bool wet_out = false;  // the cell's zeta written below is not NaN
  if (t < T) {
    const int i = c0 + t;
    const uint64_t meta = pf_meta;
    const uint64_t ln = pf_ln;
    const CellS si = ld(&g.cs[i]);
    Tri tr;
    if constexpr (MODE == 2) {
      const double4 q0 = g.vxyz[v0], q1 = g.vxyz[v1], q2 = g.vxyz[v2];
      tr.x[0] = q0.x; tr.y[0] = q0.y; tr.zv[0] = q0.z;
      tr.x[1] = q1.x; tr.y[1] = q1.y; tr.zv[1] = q1.z;
      tr.x[2] = q2.x; tr.y[2] = q2.y; tr.zv[2] = q2.z;
    } else {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        constexpr int TM = tile_tmax(CAP);
        tr.x[k] = L.vx[k * TM + t]; tr.y[k] = L.vy[k * TM + t]; tr.zv[k] = L.vz[k * TM + t];
      }
    }
    const double hi = F.h, umag_in = F.umag;
    const bool dry_i = F.dry;
    const double zi = si.z;
    const double area = tri_area(tr);
    double p0 = ZERO, p1 = ZERO, p2 = ZERO;
    if constexpr (MODE == 2) {
      if (S.phis_lo) {
        const double4 v = S.phis_lo[S.rs_pos[i]];
        p0 = v.x; p1 = v.y; p2 = v.z;
      } else {
        p0 = S.phis[i]; p1 = S.phis[N + i]; p2 = S.phis[2 * N + i];
      }
    } else {
    const double dhx = F.dhx, dhy = F.dhy;
    const double xci = si.xc, yci = si.yc;
    auto side_flux = [&](uint32_t bits, const Side& sd, int j, double& d0, double& d1,
                         double& d2) {
      const double fho = L.fh[j * CAP + t];
      double fuo = L.fu[j * CAP + t], fvo = L.fv[j * CAP + t];
      if ((bits & CMS_KIND) == SK_INTERIOR) {
        const int k = lid(ln, j), nl = (bits >> CMS_NL) & 3;
        Nbr nb;
        nb.fh = L.fh[nl * CAP + k]; nb.fu = L.fu[nl * CAP + k]; nb.fv = L.fv[nl * CAP + k];
        nb.z = L.z[k]; nb.h = L.h[k]; nb.dry = L.dry[k];
        if (XSF) {
          // invisc_fluxes_xsf.f90:65-66,79-80: both states of the edge take
          // the y-gradients of its right cell (edges%e(2))
          const int r = ((bits >> CMS_LEFT) & 1) ? k : t;
          const double gu = L.gy[r], gv = L.gy[CAP + r];
          fuo = dry_i ? ZERO : fuo + gu * L.rcy[j * CAP + t];
          fvo = dry_i ? ZERO : fvo + gv * L.rcy[j * CAP + t];
          nb.fu = nb.dry ? ZERO : nb.fu + gu * L.rcy[nl * CAP + k];
          nb.fv = nb.dry ? ZERO : nb.fv + gv * L.rcy[nl * CAP + k];
        }
        interior_flux(o, bits, sd, dry_i, hi, zi, fho, fuo, fvo, nb, d0, d1, d2);
      } else {
        if (XSF) {  // the cell's own gradient, as invisc_fluxes' boundary routines
          fuo = fuo + L.gy[t] * L.rcy[j * CAP + t];
          fvo = fvo + L.gy[CAP + t] * L.rcy[j * CAP + t];
        }
        const int va = g.cvtx[j * N + i], vb = g.cvtx[(j == 2 ? 0 : j + 1) * N + i];
        const double ha = S.hvtx[va], hb2 = S.hvtx[vb];
        const bool rev = (bits >> CMS_REV) & 1;
        boundary_flux(o, ri, S.sq, S.sh, s_at, s_null, s_skip, bits, sd, fho, fuo, fvo, hi,
                      dry_i, rev ? hb2 : ha, rev ? ha : hb2, d0, d1, d2);
      }
    };
    // the sides in the reference's accumulation order (invisc_fluxes:
    // interior edges by number, walls, boundary routines), each added to phi
    // as it is formed
    if constexpr (EDGES) {
#pragma unroll
      for (int tt = 0; tt < 3; ++tt) {
        const int j = (int)(meta >> (2 * tt)) & 3;
        const uint32_t kind = slot_bits(meta, j) & CMS_KIND;
        if (tile_wet || (kind != SK_INTERIOR && kind != SK_WALL)) {
          p0 = p0 + L.fh[j * CAP + t]; p1 = p1 + L.fu[j * CAP + t]; p2 = p2 + L.fv[j * CAP + t];
        } else {
          p0 = p0 + ZERO; p1 = p1 + ZERO; p2 = p2 + ZERO;
        }
      }
    } else {
#pragma unroll 1  // one side at a time: fewer live registers
    for (int tt = 0; tt < 3; ++tt) {
      const int j = (int)(meta >> (2 * tt)) & 3;
      const uint32_t bits = slot_bits(meta, j), kind = bits & CMS_KIND;
      // no flux between two dry cells (invisc_fluxes.f90:47) nor along a dry
      // cell's wall (wenslp3_fluxes.f90:40, weslp2_fluxes.f90:40): the side's
      // geometry is not formed
      const bool none = dry_i && (kind == SK_WALL || (kind == SK_INTERIOR && L.dry[lid(ln, j)]));
      double d0 = ZERO, d1 = ZERO, d2 = ZERO;
      if (!none) side_flux(bits, side_dyn<CAP>(L, t, j), j, d0, d1, d2);
      p0 = p0 + d0; p1 = p1 + d1; p2 = p2 + d2;
    }
    }  // EDGES
    if (o.friction == 1) {
      // fric_terms.f90:24-98 (explicit Manning), after the edge fluxes
      double r = dry_i ? ZERO : G * si.cd * si.cd / pow_one_third(hi);
      r = -r * umag_in;
      p1 = p1 + r * F.up * area;
      p2 = p2 + r * F.vp * area;
    }
    if (!dry_i) {
      // bed_slope_STD.f90:23-70 with adjustZb.f90:41-70 and bed_slope_VTX.f90:32-33
      double hp[3];
      bool adjust = false;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        hp[k] = hi + dhx * (tr.x[k] - xci) + dhy * (tr.y[k] - yci);
        if (hp[k] < ZERO) adjust = true;
      }
      double gx, gy;
      if (adjust) {
        double zp[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          hp[k] = dmin(ZERO, hp[k]);
          zp[k] = tr.zv[k] + hp[k];
        }
        const double bb0 = tr.y[1] - tr.y[2], cc0 = tr.x[2] - tr.x[1];
        const double bb1 = tr.y[2] - tr.y[0], cc1 = tr.x[0] - tr.x[2];
        const double bb2 = tr.y[0] - tr.y[1], cc2 = tr.x[1] - tr.x[0];
        const double a = HALF / area;
        gx = a * (bb0 * zp[0] + bb1 * zp[1] + bb2 * zp[2]);
        gy = a * (cc0 * zp[0] + cc1 * zp[1] + cc2 * zp[2]);
      } else {
        const double ngx = HALF * (tr.zv[0] * (tr.y[1] - tr.y[2]) + tr.zv[1] * (tr.y[2] - tr.y[0]) +
                                   tr.zv[2] * (tr.y[0] - tr.y[1]));
        const double ngy = HALF * (tr.zv[0] * (tr.x[2] - tr.x[1]) + tr.zv[1] * (tr.x[0] - tr.x[2]) +
                                   tr.zv[2] * (tr.x[1] - tr.x[0]));
        const Recip ra = recip(area, ngx);
        gx = qdiv(ra, ngx, area);
        gy = qdiv(ra, ngy, area);
      }
      p1 = p1 - G * hi * gx * area;
      p2 = p2 - G * hi * gy * area;
    }
    if (o.wind && !dry_i) {  // solverFVT.f90:218-224
      const double2 w = g.wind[i];
      p1 = p1 + w.x;
      p2 = p2 + w.y;
    }
    if (o.ncv && ((meta >> CM_CULVERT) & 1)) {  // solverFVT.f90:227-233, culvert order
      for (int m = 0; m < o.ncv; ++m) {
        const int2 io = g.cv_io[m];
        if (io.x == i) p0 = p0 - g.cvsrc[m];
        if (io.y == i) p0 = p0 + g.cvsrc[m];
      }
    }
    if constexpr (MODE == 1) {
      S.phi0[i] = p0; S.phi0[N + i] = p1; S.phi0[2 * N + i] = p2;
    }
    }  // MODE != 2
    if constexpr (MODE != 1) {
    const double perim = side<0>(tr).len + side<1>(tr).len + side<2>(tr).len;
    // solverFVT.f90:259-283
    double delt = c->dt;
    if (o.steady) {
      const double wc = dry_i ? ZERO : sqrt(G * hi);
      double d = umag_in + wc;
      if (d < SMALL) d = 1.044;
      delt = o.cfl * area / (d * perim);
    }
    const double dt = delt / area;
    // the SSP combination: slots 0 and s-1 in ascending order (the slots in
    // between have alpha 0 in every table: rk_table)
    double qh = ZERO, qu = ZERO, qv = ZERO;
    // the cell's rk slots 0 and s-1, the only ones with a nonzero alpha in
    // the SSP tables; hgradfix2 zeroed the hu, hv of slot s-1 above
    const double* s0 = st.slot[0];
    double rh0 = s0[i], ru0 = s0[N + i], rv0 = s0[2 * N + i];
    double rh1 = ZERO, ru1 = ZERO, rv1 = ZERO;
    if (st.s >= 2) {
      const double* s1 = st.slot[st.s - 1];
      rh1 = s1[i]; ru1 = s1[N + i]; rv1 = s1[2 * N + i];
    }
    if (MODE != 2 && F.fix) {
      if (st.s >= 2) { ru1 = ZERO; rv1 = ZERO; }
      else { ru0 = ZERO; rv0 = ZERO; }
    }
    if (st.alpha[0] != ZERO) {
      const double a = st.alpha[0];
      qh = qh + a * rh0; qu = qu + a * ru0; qv = qv + a * rv0;
    }
    if (st.s >= 2 && st.alpha[st.s - 1] != ZERO) {
      const double a = st.alpha[st.s - 1];
      qh = qh + a * rh1; qu = qu + a * ru1; qv = qv + a * rv1;
    }
    const double h0n = rh0;  // h at the start of the step (residual)
    qh = qh + st.beta * dt * p0;
    qu = (qu + st.beta * dt * p1);
    qv = (qv + st.beta * dt * p2);
    if (o.vclip) {
      // solverFVT.f90:286-296
      const double heff = dmax(h_dry, qh);
      const double um = sqrt(qu * qu + qv * qv) / heff;
      if (um > o.vceiling) { qu = o.vceiling * qu / um; qv = o.vceiling * qv / um; }
    }
    // solverFVT.f90:300-325
    qh = dmax(qh, ZERO);
    const double hn = qh;
    const double hd = dry_i ? h_wet : h_dry;
    double un, vn;
    if (hn > hd) { un = qu / hn; vn = qv / hn; }
    else { qu = ZERO; qv = ZERO; un = ZERO; vn = ZERO; }
    if (!st.last) {
      st.out[i] = qh;
      st.out[N + i] = qu;
      st.out[2 * N + i] = qv;
      double um;
      const bool dn = find_dry(hn, un, vn, h_dry, h_wet, &um);
      st4(&S.dout[i], hn, un, vn, dn ? ONE : ZERO);
      S.zeta[i] = dn ? __builtin_nan("") : hn + zi;
      wet_out = !dn;
    } else {
      // find_dry_cells + fric_terms_impl (solverFVT.f90:335-336)
      double um;
      const bool dpost = find_dry(hn, un, vn, h_dry, h_wet, &um);
      S.dry_post[i] = dpost;
      double rough = ZERO;
      if (!dpost) {
        const double fdt = o.steady ? s_dlast : c->dt;
        rough = G * si.cd * si.cd / pow_one_third(hn);
        rough = hn / (hn + fdt * rough * um);
      }
      if (o.friction == 0) {
        un = un * rough;
        vn = vn * rough;
        if (!dpost) { qu = hn * un; qv = hn * vn; }
      }
      // residual terms (solverFVT.f90:341-346)
      const double dh = hn - h0n;
      top = dh * dh;
      bot = h0n * h0n;
      if (o.stop_cr == 2) {  // discharge terms of stopping test 2
        const double qu0 = st.slot[0][N + i], qv0 = st.slot[0][2 * N + i];
        const double du = qu - qu0, dv = qv - qv0;
        topq = du * du + dv * dv;
        botq = qu0 * qu0 + qv0 * qv0;
      }
      // hvalve (solverFVT.f90:357-370): outward normal of the cell's side
      if (o.hvalve) {
        if ((slot_bits(meta, 0) >> CMS_HBAR) & 1) {
          const Side sd = side<0>(tr);
          if (un * sd.nx + vn * sd.ny < 0) { un = ZERO; vn = ZERO; }
        }
        if ((slot_bits(meta, 1) >> CMS_HBAR) & 1) {
          const Side sd = side<1>(tr);
          if (un * sd.nx + vn * sd.ny < 0) { un = ZERO; vn = ZERO; }
        }
        if ((slot_bits(meta, 2) >> CMS_HBAR) & 1) {
          const Side sd = side<2>(tr);
          if (un * sd.nx + vn * sd.ny < 0) { un = ZERO; vn = ZERO; }
        }
      }
      st.out[i] = qh;
      st.out[N + i] = qu;
      st.out[2 * N + i] = qv;
      // phi reset at dry cells (solverFVT.f90:376-386)
      if (dpost) { p0 = ZERO; p1 = ZERO; p2 = ZERO; }
      S.phi[i] = p0; S.phi[N + i] = p1; S.phi[2 * N + i] = p2;
      mxphi = dmax(dmax(dmax(mxphi, fabs(p0)), fabs(p1)), fabs(p2));
      // stage-1 dry flags and CFL candidate of the next step
      double um2;
      const bool dn = find_dry(hn, un, vn, h_dry, h_wet, &um2);
      st4(&S.dout[i], hn, un, vn, dn ? ONE : ZERO);
      S.zeta[i] = dn ? __builtin_nan("") : hn + zi;
      wet_out = !dn;
      const double wc = dn ? ZERO : sqrt(G * hn);
      cand = o.cfl * area / ((um2 + wc + SMALL) * perim);
    }
    }  // MODE != 1
  }
  if (MODE != 1) {
    // the tile's flag for the next vertex pass, written with its zeta
    const int any_wet = __syncthreads_or(wet_out);
    if (t == 0) S.twet[b] = any_wet != 0;
  }
  if (MODE == 1 || !st.last) return;
  // per-tile partials: a fixed butterfly within each wave (every lane ends
  // with the same bits), then the waves' results in wave order
  const bool q2 = o.stop_cr == 2;
  double r0 = top, r1 = bot, r2 = mxphi, r3 = cand, r4 = topq, r5 = botq;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    r0 = r0 + __shfl_xor(r0, off);
    r1 = r1 + __shfl_xor(r1, off);
    r2 = dmax(r2, __shfl_xor(r2, off));
    r3 = dmin(r3, __shfl_xor(r3, off));
    if (q2) {
      r4 = r4 + __shfl_xor(r4, off);
      r5 = r5 + __shfl_xor(r5, off);
    }
  }
  constexpr int NW = CAP / 64;
  __shared__ double s_red[NW][PART];
  if ((t & 63) == 0) {
    double* q = s_red[t >> 6];
    q[0] = r0; q[1] = r1; q[2] = r2; q[3] = r3; q[4] = r4; q[5] = r5;
  }
  __syncthreads();
  if (t == 0) {
    double a0 = s_red[0][0], a1 = s_red[0][1], a2 = s_red[0][2], a3 = s_red[0][3];
    double a4 = s_red[0][4], a5 = s_red[0][5];
#pragma unroll
    for (int w = 1; w < NW; ++w) {
      a0 = a0 + s_red[w][0];
      a1 = a1 + s_red[w][1];
      a2 = dmax(a2, s_red[w][2]);
      a3 = dmin(a3, s_red[w][3]);
      a4 = a4 + s_red[w][4];
      a5 = a5 + s_red[w][5];
    }
    S.part[PART * b] = a0;
    S.part[PART * b + 1] = a1;
    S.part[PART * b + 2] = a2;
    S.part[PART * b + 3] = a3;
    S.part[PART * b + 4] = q2 ? a4 : ZERO;
    S.part[PART * b + 5] = q2 ? a5 : ZERO;
  }
}

// to_vertices' cell pass alone (solverFVT.f90:483 -> to_vertices.f90:44-55),
// after k_vertex, for the final state
__global__ void k_final_cells(Geo g, const uint8_t* dry, const uint8_t* vwet, CellD* d) {
  const int i = xcd_tile() * blockDim.x + threadIdx.x;
  if (i >= g.N) return;
  if (dry ? (dry[i] != 0) : (d[i].dry != ZERO)) return;
  const int N = g.N;
  const bool flag = vwet[g.cvtx[i]] && vwet[g.cvtx[N + i]] && vwet[g.cvtx[2 * N + i]];
  if (!flag) { d[i].u = ZERO; d[i].v = ZERO; }
}

// The depth gradient delh as the last stage left it (cell_grads.f90:24-75 then
// hgradfix2.f90:46-98 on that stage's input), which error_est reads after the
// loop (solverFVT.f90:487, error_est.f90:80-82).  The stage kernel keeps it in
// registers; this pass forms it again from the same input with the same code.
// dh: [2N], x components then y components.
__global__ void k_delh(Geo g, Opts o, const CellD* din, const double* gwork, double* dh) {
  const int i = xcd_tile() * blockDim.x + threadIdx.x;
  if (i >= g.N) return;
  const int N = g.N;
  const Tri tr = load_tri(g, i);
  Faces F;
  cell_faces(g, o, din, gwork, i, g.cmeta[i], g.cnbr[i], g.cnbr[N + i], g.cnbr[2 * N + i], tr,
             true, F);
  dh[i] = F.dhx;
  dh[N + i] = F.dhy;
}

// delh as k_delh forms it, written in input cell order (dh[perm[i]],
// dh[N + perm[i]]) so that it can be copied to the caller as it is.
__global__ void k_delh_input_order(Geo g, Opts o, const CellD* din, const double* gwork,
                                   const int* perm, double* dh) {
  const int i = xcd_tile() * blockDim.x + threadIdx.x;
  if (i >= g.N) return;
  const int N = g.N;
  const Tri tr = load_tri(g, i);
  Faces F;
  cell_faces(g, o, din, gwork, i, g.cmeta[i], g.cnbr[i], g.cnbr[N + i], g.cnbr[2 * N + i], tr,
             true, F);
  const int j = perm[i];
  dh[j] = F.dhx;
  dh[N + j] = F.dhy;
}

// ---------------------------------------------------------------- culverts
// Culvert tables: per culvert the inlet's offset from its cell's centroid,
// the invert elevation, and the performance curve head[tptr[m] ..
// tptr[m+1]) -> q.
struct CvTab {
  const double2* rin;
  const double* zinv;
  const int* tptr;
  const double *head, *q;
};

// cvdisch.f90:26-58, one thread per culvert, before the stage kernel: the
// water surface at the inlet point reconstructed from the inlet cell's
// stage input with that stage's depth gradient (delh after cell_grads and
// hgradfix2, formed by the stage kernel's own code) plus the bed gradient of
// bed_slope_VTX.f90:32-33, so delh + delz as the reference adds them.
__global__ void k_cv_src(Geo g, Opts o, const CellD* din, const double* gwork, CvTab cv,
                         double* cvsrc) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= o.ncv) return;
  const int N = g.N;
  const int i = g.cv_io[m].x;
  const CellD di = ld(&din[i]);
  double src = ZERO;
  if (di.dry == ZERO) {
    const Tri tr = load_tri(g, i);
    Faces F;
    cell_faces(g, o, din, gwork, i, g.cmeta[i], g.cnbr[i], g.cnbr[N + i], g.cnbr[2 * N + i], tr,
               true, F);
    const double area = tri_area(tr);
    const double nzx = HALF * (tr.zv[0] * (tr.y[1] - tr.y[2]) + tr.zv[1] * (tr.y[2] - tr.y[0]) +
                               tr.zv[2] * (tr.y[0] - tr.y[1]));
    const double nzy = HALF * (tr.zv[0] * (tr.x[2] - tr.x[1]) + tr.zv[1] * (tr.x[0] - tr.x[2]) +
                               tr.zv[2] * (tr.x[1] - tr.x[0]));
    const Recip ra = recip(area, nzx);
    const double dzx = qdiv(ra, nzx, area), dzy = qdiv(ra, nzy, area);
    const double2 r = cv.rin[m];
    const double zeta = di.h + g.cs[i].z;
    double head = zeta + (F.dhx + dzx) * r.x + (F.dhy + dzy) * r.y;
    head = head - cv.zinv[m];
    if (!(head < o.h_dry)) {
      const int a = cv.tptr[m], k = cv.tptr[m + 1] - a;
      const double *hh = cv.head + a, *qq = cv.q + a;
      if (head < hh[0]) {
        src = qq[0];
      } else if (head >= hh[k - 1]) {
        src = qq[k - 1];
      } else {
        for (int j = 0; j < k - 1; ++j)
          if (head < hh[j + 1]) {
            const double tt = (head - hh[j]) / (hh[j + 1] - hh[j]);
            src = qq[j + 1] * tt + qq[j] * (ONE - tt);
            break;
          }
      }
    }
  }
  cvsrc[m] = src;
}

// ---------------------------------------------------------------- residual smoothing
// The level schedule of the Gauss-Seidel sweep (see rsavg_schedule).
struct RsTab {
  const int *lptr, *cell, *nptr, *nbr, *rlev;
};
constexpr int RS_THREADS = 256;

// solverFVT.f90:238-256: nsweep sweeps of phi(i) = fac1 phi0(i) + fac2 phi(j)
// over i's t2t row in row order, one block per reach walking its levels.
// A level's cells read cells of lower levels (already written in this
// sweep) and of higher levels (still holding the previous sweep, or phi0 in
// the first), so one barrier per level gives the serial sweep's values.
__global__ void __launch_bounds__(RS_THREADS)
k_rsavg(RsTab rs, const Ctrl* ctrl, int N, int nsweep, double fac1, double fac2,
        const double* phi0, double* phis) {
  const int r = blockIdx.x;
  if (ctrl[r].done) return;
  const int l0 = rs.rlev[r], l1 = rs.rlev[r + 1];
  for (int sw = 0; sw < nsweep; ++sw) {
    const bool first = sw == 0;
    for (int l = l0; l < l1; ++l) {
      const int k1 = rs.lptr[l + 1];
      for (int k = rs.lptr[l] + (int)threadIdx.x; k < k1; k += RS_THREADS) {
        const int i = rs.cell[k];
        double q0 = fac1 * phi0[i], q1 = fac1 * phi0[N + i], q2 = fac1 * phi0[2 * N + i];
        const int m1 = rs.nptr[k + 1];
        for (int m = rs.nptr[k]; m < m1; ++m) {
          const int e = rs.nbr[m];
          const int j = e & 0x7fffffff;
          const double* src = (first && e < 0) ? phi0 : phis;
          q0 = q0 + fac2 * src[j];
          q1 = q1 + fac2 * src[N + j];
          q2 = q2 + fac2 * src[2 * N + j];
        }
        phis[i] = q0; phis[N + i] = q1; phis[2 * N + i] = q2;
      }
      __syncthreads();
    }
  }
}

// The same sweep with every value a level reads kept in shared memory (the
// tables of Prep::rs_fast): a ring of RSF_W position slots, each holding a
// position's previous-sweep value from RSF_G + 1 levels before its own level
// and its new value after it, plus RSF_FS far slots for values read outside
// that window.  A level then waits only on shared memory (its barrier fences
// shared memory alone), and the global loads of the levels ahead (records,
// phi0, previous-sweep values) are issued RSF_D levels early.  Positions are
// in level order; the sweep reads and writes level-ordered copies (one
// double4 per position, w unused) and also writes the cell-ordered result
// the update pass reads.  Records are packed so that a level issues 9 loads
// and 2 stores: the RSF_D - 1 levels in flight stay within the 6-bit vector
// memory counter, whose waits are then exact, and the block's one compute
// unit moves few cache lines per level (its memory pipeline, not latency,
// bounds a level).
struct RsFast {
  const int *rlev, *pptr, *psrc, *pdst;
  const int2* cf;    // per position {cell, far slot it writes or -1}
  const int4* code;  // per position its RSF_E entries (Prep::rs_code) as 16-bit integers
  const int4* hdr;   // per level two int4 (Prep::rs_hdr); a zero one at index L
  int N, NF, L;      // positions, old far values filled by levels, levels
};
constexpr int RSF_D = 4;
// bytes of the ring, the far slots and one dummy slot
constexpr int RSF_LDS = 3 * 8 * (RSF_W + RSF_FS + 1);

// phi0 in level order (lo[k] = phi0 of cell[k])
__global__ void k_rs_order(RsFast T, const double* phi0, double4* lo) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= T.N) return;
  const int i = T.cf[k].x, N = T.N;
  lo[k] = make_double4(phi0[i], phi0[N + i], phi0[2 * N + i], ZERO);
}

// the previous sweep's values of the old far values levels fill, with their
// far slot in w (fval[f] = src_lo[fsrc[f]], w = fdst[f])
__global__ void k_rs_far(RsFast T, const int* fsrc, const int* fdst, const double4* src_lo,
                         double4* fval) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= T.NF) return;
  const double4 v = src_lo[fsrc[f]];
  fval[f] = make_double4(v.x, v.y, v.z, (double)fdst[f]);
}

// Registers of one pipeline stage: the records of level l and the header
// of level l + RSF_D.  Lanes without work load from valid addresses and
// write to the dummy slot, so the loop body is straight-line code.
struct RsSet {
  int k, cell, farw, q;
  bool vk, vq, vf;  // the lane has a position; a ring fill; a far fill
  int4 c;           // the entries, 16 bits each
  double4 p, v, g;  // phi0; previous value of the fill position; far fill (slot in w)
  int4 h0, h1;      // header of level l + RSF_D
};

__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The header of level l (a zero one past the reach), read by vector loads:
// a scalar load would hold up every level's barrier, which waits for the
// scalar memory counter it shares with shared memory.
__device__ __forceinline__ void rs_header(const RsFast& T, int l, int l1, int4& h0, int4& h1) {
  int idx = l < l1 ? l : T.L;
  asm volatile("v_mov_b32 %0, %1" : "=v"(idx) : "v"(idx));
  h0 = T.hdr[2 * idx];
  h1 = T.hdr[2 * idx + 1];
}

// Loads of level l (whose header s.h0/s.h1 holds) and the header of level
// l + RSF_D: position t's record and phi0, the previous value of fill
// position t, and old far value t.
__device__ __forceinline__ void rs_load(RsSet& s, const RsFast& T, const double4* p0lo,
                                        const double4* src_lo, const double4* fval, int l,
                                        int l1, int t) {
  const int4 h0 = s.h0, h1 = s.h1;  // {position start, width, fill start, fill width}, {far start, far count}
  s.vk = t < h0.y;
  const int k = h0.x + (s.vk ? t : 0);
  s.k = k;
  const int2 cf = T.cf[k];
  s.cell = cf.x;
  s.farw = cf.y;
  s.c = T.code[k];
  s.p = p0lo[k];
  s.vq = t < h0.w;
  const int q = h0.z + (s.vq ? t : 0);
  s.q = q;
  s.v = src_lo[q];
  s.vf = t < h1.y;
  s.g = fval[s.vf ? h1.x + t : 0];
  rs_header(T, l + RSF_D, l1, s.h0, s.h1);
}

// One level: position t's smoothed value from shared memory, written to its
// ring slot (and far slot) and to global memory; then the fills.
__device__ __forceinline__ void rs_level(const RsSet& s, double* sm, int base, double fac1,
                                         double fac2, double4* out_lo, double* dummy) {
  const int dum = RSF_W + RSF_FS;
  const int code[RSF_E] = {s.c.x << 16 >> 16, s.c.x >> 16, s.c.y << 16 >> 16, s.c.y >> 16,
                           s.c.z << 16 >> 16, s.c.z >> 16, s.c.w << 16 >> 16, s.c.w >> 16};
  // every entry's slot read at once (the dummy slot past a row's end);
  // past the end the term is -0.0, which leaves any sum unchanged.  Rows
  // of more than four entries (boundary cells) sit last in their level, so
  // most waves skip the second half.
  double q0 = fac1 * s.p.x, q1 = fac1 * s.p.y, q2 = fac1 * s.p.z;
  auto half = [&](int e0) {
    double a[4][3];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double* v = sm + 3 * (code[e0 + e] < 0 ? dum : code[e0 + e]);
      a[e][0] = v[0]; a[e][1] = v[1]; a[e][2] = v[2];
    }
    // (empty statements that take the values: the reads stay unconditional,
    // issued together, rather than one branch and wait per entry)
#pragma unroll
    for (int e = 0; e < 4; ++e) asm volatile("" : "+v"(a[e][0]), "+v"(a[e][1]), "+v"(a[e][2]));
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const bool end = code[e0 + e] < 0;
      q0 = q0 + (end ? -ZERO : fac2 * a[e][0]);
      q1 = q1 + (end ? -ZERO : fac2 * a[e][1]);
      q2 = q2 + (end ? -ZERO : fac2 * a[e][2]);
    }
  };
  half(0);
  if (code[4] >= 0) half(4);
  double* w = sm + 3 * (s.vk ? ((s.k - base) & (RSF_W - 1)) : dum);
  w[0] = q0; w[1] = q1; w[2] = q2;
  double* f = sm + 3 * ((s.vk && s.farw >= 0) ? RSF_W + s.farw : dum);
  f[0] = q0; f[1] = q1; f[2] = q2;
  double4* o = s.vk ? out_lo + s.k : reinterpret_cast<double4*>(dummy);
  *o = make_double4(q0, q1, q2, ZERO);
  double* r = sm + 3 * (s.vq ? ((s.q - base) & (RSF_W - 1)) : dum);
  r[0] = s.v.x; r[1] = s.v.y; r[2] = s.v.z;
  double* g = sm + 3 * (s.vf ? RSF_W + (int)s.g.w : dum);
  g[0] = s.g.x; g[1] = s.g.y; g[2] = s.g.z;
}

// One sweep of every reach (a block each): src_lo holds the previous
// sweep's values in level order (phi0's in the first sweep).  dummy: four
// doubles that lanes without a position write to.
__global__ void __launch_bounds__(RSF_NT)
k_rsavg_fast(RsFast T, const Ctrl* ctrl, double fac1, double fac2, const double4* p0lo,
             const double4* src_lo, const double4* fval, double4* out_lo, double* dummy) {
  extern __shared__ double sm[];
  const int r = blockIdx.x;
  if (ctrl[r].done) return;
  const int t = threadIdx.x;
  const int l0 = T.rlev[r], l1 = T.rlev[r + 1];
  const int base = T.hdr[2 * l0].x;
  // prologue: the ring slots of the first RSF_G + 1 levels and the far slots
  // read before any level fills them
  for (int m = T.pptr[r] + t; m < T.pptr[r + 1]; m += (int)blockDim.x) {
    const double4 v = src_lo[T.psrc[m]];
    double* w = sm + 3 * T.pdst[m];
    w[0] = v.x; w[1] = v.y; w[2] = v.z;
  }
  RsSet s0, s1, s2, s3;
  rs_header(T, l0, l1, s0.h0, s0.h1);
  rs_header(T, l0 + 1, l1, s1.h0, s1.h1);
  rs_header(T, l0 + 2, l1, s2.h0, s2.h1);
  rs_header(T, l0 + 3, l1, s3.h0, s3.h1);
  rs_load(s0, T, p0lo, src_lo, fval, l0, l1, t);
  rs_load(s1, T, p0lo, src_lo, fval, l0 + 1, l1, t);
  rs_load(s2, T, p0lo, src_lo, fval, l0 + 2, l1, t);
  rs_load(s3, T, p0lo, src_lo, fval, l0 + 3, l1, t);
  __syncthreads();
  for (int l = l0; l < l1; l += RSF_D) {
    rs_level(s0, sm, base, fac1, fac2, out_lo, dummy);
    rs_load(s0, T, p0lo, src_lo, fval, l + RSF_D, l1, t);
    lds_barrier();
    if (l + 1 >= l1) break;
    rs_level(s1, sm, base, fac1, fac2, out_lo, dummy);
    rs_load(s1, T, p0lo, src_lo, fval, l + 1 + RSF_D, l1, t);
    lds_barrier();
    if (l + 2 >= l1) break;
    rs_level(s2, sm, base, fac1, fac2, out_lo, dummy);
    rs_load(s2, T, p0lo, src_lo, fval, l + 2 + RSF_D, l1, t);
    lds_barrier();
    if (l + 3 >= l1) break;
    rs_level(s3, sm, base, fac1, fac2, out_lo, dummy);
    rs_load(s3, T, p0lo, src_lo, fval, l + 3 + RSF_D, l1, t);
    lds_barrier();
  }
}

// ---------------------------------------------------------------- output fields
// Tables of the final to_vertices' velocity part (enforce_bcs.f90): per
// vertex the outflow-string normals its velocity is checked against and its
// wall tangent; per inflow string its nodes, node normals and edge lengths.
struct OutTab {
  const int *vo_ptr, *vwall;
  const double2 *vo_nrm, *wtang;
  const int *is_ptr, *is_node;
  const double2* is_pn;
  const double* is_len;
};

// to_vertices.f90:56-65 after the cell pass, for vertex p: the vertex
// velocities and drag coefficient by ctr2vtx.f90:73-80 (type 3: the
// inverse-distance weights of c2v_setup.f90:106-119 normalised by their sum,
// over every incident cell in n2t order), wetdryvfix.f90:30-49 (no-slip
// walls), zero velocity where the interpolated depth is below h_dry, then
// the vertex part of enforce_bcs.f90:30-106: the stage strings' depths, the
// outflow strings' backflow test, the walls and the dry vertices.  hvtx holds
// the interpolated depth on entry (k_vertex with no strings) and the final
// depth on exit.  The inflow strings are k_out_inflow's.
__global__ void __launch_bounds__(256)
k_out_vertex(Geo g, Opts o, const CellD* d, const uint8_t* dry, OutTab T, const double* sh,
             double* hvtx, double* uvtx, double* vvtx, double* cdvtx) {
  const int p = xcd_tile() * blockDim.x + threadIdx.x;
  if (p >= g.P) return;
  const double4 q = g.vxyz[p];
  const int t0 = g.vptr[p], t1 = g.vptr[p + 1];
  double tdist = ZERO;
  for (int t = t0; t < t1; ++t) {
    const CellS sk = ld(&g.cs[g.vcell[t]]);
    const double xk = sk.xc - q.x, yk = sk.yc - q.y;
    tdist = tdist + ONE / sqrt(xk * xk + yk * yk);
  }
  double a = ZERO, b = ZERO, c = ZERO;
  bool wett = false, dryt = false;
  for (int t = t0; t < t1; ++t) {
    const int k = g.vcell[t];
    const CellS sk = ld(&g.cs[k]);
    const CellD dk = ld(&d[k]);
    const double xk = sk.xc - q.x, yk = sk.yc - q.y;
    const double w = (ONE / sqrt(xk * xk + yk * yk)) / tdist;
    a = a + w * dk.u;
    b = b + w * dk.v;
    c = c + w * sk.cd;
    if (dry ? (dry[k] != 0) : (dk.dry != ZERO)) dryt = true; else wett = true;
  }
  double u = a, v = b;
  if (o.btype == 0 && wett && dryt) { u = ZERO; v = ZERO; }
  const double hpre = hvtx[p];
  if (hpre < o.h_dry) { u = ZERO; v = ZERO; }
  const int vb = g.vbc[p];
  const double h = (vb >= 0) ? dmax(ZERO, sh[vb] - q.z) : hpre;
  for (int e = T.vo_ptr[p]; e < T.vo_ptr[p + 1]; ++e) {
    const double2 n = T.vo_nrm[e];
    if (u * n.x + v * n.y < ZERO) { u = ZERO; v = ZERO; }
  }
  const int w = T.vwall[p];
  if (w >= 0) {
    if (o.btype == 0) { u = ZERO; v = ZERO; }
    else {
      const double2 tg = T.wtang[w];
      const double dot = u * tg.x + v * tg.y;
      u = dot * tg.x;
      v = dot * tg.y;
    }
  }
  if (h < o.h_dry) { u = ZERO; v = ZERO; }
  hvtx[p] = h;
  uvtx[p] = u;
  vvtx[p] = v;
  cdvtx[p] = c;
}

// enforce_bcs.f90:117-149 with vbc_by_h.f90:32-80: the velocities of the
// inflow strings' nodes (vtype 0: the discharge value along the node normal;
// vtype 1 and 5: the node speeds that carry the string's discharge on the
// vertex depths), one block per reach, the reach's strings in order (a node
// on two strings keeps the later one's value).  vtype 1/5 strings have at
// most QMAX edges; their speeds are formed by thread 0 in string order.
__global__ void __launch_bounds__(256)
k_out_inflow(Geo g, Opts o, OutTab T, const double* sq, const double* hvtx, double* uvtx,
             double* vvtx) {
  __shared__ double s_v[QMAX + 1], s_u[QMAX + 1], s_a[QMAX + 1];
  const RInfo ri = g.reach[blockIdx.x];
  const int t = threadIdx.x;
  const double h_dry = o.h_dry;
  for (int ch = ri.in_base; ch < ri.in_base + ri.n_in; ++ch) {
    const int vt = g.in_type[ch];
    if (vt != 0 && vt != 1 && vt != 5) continue;
    const int a0 = T.is_ptr[ch], n = T.is_ptr[ch + 1] - a0;
    const double qin = sq[ch];
    if (vt != 0 && t == 0) {
      // vbc_by_h.f90:45-73
      double ht = ZERO, at = ZERO;
      for (int i = 0; i + 1 < n; ++i) {
        const double h1 = hvtx[T.is_node[a0 + i]], h2 = hvtx[T.is_node[a0 + i + 1]];
        s_a[i] = T.is_len[a0 + i] * (h1 + h2) * HALF;
        ht = ht + (h1 + h2) * HALF;
        at = at + s_a[i];
      }
      for (int i = 0; i + 1 < n; ++i) {
        const double h1 = hvtx[T.is_node[a0 + i]], h2 = hvtx[T.is_node[a0 + i + 1]];
        s_u[i] = HALF * (h1 + h2) * qin / ht / at;
      }
      for (int i = 0; i < n; ++i) s_v[i] = ZERO;
      for (int i = 1; i + 1 < n; ++i)
        if (hvtx[T.is_node[a0 + i]] > h_dry) s_v[i] = HALF * (s_u[i - 1] + s_u[i]);
      for (int i = 0; i + 1 < n; ++i) s_u[i] = HALF * (s_v[i] + s_v[i + 1]);
      double qq = ZERO;
      for (int i = 0; i + 1 < n; ++i) qq = qq + s_u[i] * s_a[i];
      for (int i = 1; i + 1 < n; ++i)
        if (hvtx[T.is_node[a0 + i]] > h_dry) s_v[i] = s_v[i] * qin / qq;
    }
    __syncthreads();
    for (int i = t; i < n; i += blockDim.x) {
      const int p = T.is_node[a0 + i];
      if (hvtx[p] > h_dry) {
        const double sp = (vt == 0) ? qin : s_v[i];
        const double2 pn = T.is_pn[a0 + i];
        uvtx[p] = sp * pn.x;
        vvtx[p] = sp * pn.y;
      }
    }
    __syncthreads();
  }
}

// error_est.f90:57-84 (etype 2), for cell i: the water-surface gradient (the
// bed's P1 gradient plus delh, which cell_grads and hgradfix2 form from the
// last stage's input) times a third of the perimeter over the water-surface
// elevation, on cells with h >= h_dry.  dh_in: that stage's input (grad 0:
// no stage has run); dfin: the final state.
__global__ void k_out_cells(Geo g, Opts o, const CellD* dh_in, const double* gwork,
                            const CellD* dfin, int grad, double* e_est) {
  const int i = xcd_tile() * blockDim.x + threadIdx.x;
  if (i >= g.N) return;
  const int N = g.N;
  const double h = dfin[i].h;
  if (h < o.h_dry) { e_est[i] = ZERO; return; }
  const Tri tr = load_tri(g, i);
  Faces F;
  F.dhx = ZERO;  // before the first stage delh is zero (solverFVT.f90:63)
  F.dhy = ZERO;
  if (grad)
    cell_faces(g, o, dh_in, gwork, i, g.cmeta[i], g.cnbr[i], g.cnbr[N + i], g.cnbr[2 * N + i],
               tr, true, F);
  const double area = tri_area(tr);
  const double gzx = (tr.zv[0] * (tr.y[1] - tr.y[2]) + tr.zv[1] * (tr.y[2] - tr.y[0]) +
                      tr.zv[2] * (tr.y[0] - tr.y[1])) * HALF / area;
  const double gzy = (tr.zv[0] * (tr.x[2] - tr.x[1]) + tr.zv[1] * (tr.x[0] - tr.x[2]) +
                      tr.zv[2] * (tr.x[1] - tr.x[0])) * HALF / area;
  const double perim = side<0>(tr).len + side<1>(tr).len + side<2>(tr).len;
  const double l = perim * ONE_THIRD;
  const double gx = gzx + F.dhx, gy = gzy + F.dhy;
  e_est[i] = sqrt(gx * gx + gy * gy) * l / (g.cs[i].z + h);
}

// ---------------------------------------------------------------- RDS node routines
// The node-based routines of the reference's RDS solver on the context's
// mesh, for a bed z and depth h given at the vertices (internal order).

// part_dry.f90:30-43 and drynwet.f90:28-50 (zadjust on) for triangle i:
// the bed of its vertices, lowered at the dry vertices of a partially dry
// triangle to the water surface of its wet vertices (Brufau and
// Garcia-Navarro 2003).  With one dry vertex, its two others are
// MAX(1,MOD(i+1,4)) and MAX(1,MOD(i+2,4)) in the reference's 1-based
// numbering, which for i = 3 are both vertex 1; kept as written.  Written
// in input cell order: zout[3*perm[i] + m], pd[4*perm[i] + ...] = part_dry's
// flag, n, i, j (1-based; 0 where unset).
__global__ void k_drynwet(Geo g, double h_dry, const double* z, const double* h,
                          const int* perm, double* zout, int* pd) {
  const int i = xcd_tile() * blockDim.x + threadIdx.x;
  if (i >= g.N) return;
  const int N = g.N;
  const int v[3] = {g.cvtx[i], g.cvtx[N + i], g.cvtx[2 * N + i]};
  double zo[3] = {z[v[0]], z[v[1]], z[v[2]]};
  int n = 0, fi = 0, fj = 0;
#pragma unroll
  for (int m = 1; m <= 3; ++m)
    if (h[v[m - 1]] < h_dry) { n = n + 1; fj = m; if (n == 1) fi = m; }
  const bool part = n != 0 && n != 3;
  if (part && n == 1) {
    const int jj = max(1, (fi + 1) % 4), kk = max(1, (fi + 2) % 4);
    const double dj = z[v[jj - 1]] + h[v[jj - 1]], dk = z[v[kk - 1]] + h[v[kk - 1]];
    const double dm = dmax(dj, dk);
    zo[fi - 1] = dmin(dm, z[v[fi - 1]]);
  } else if (part) {
    const int kk = 6 - fi - fj;
    const double dk = z[v[kk - 1]] + h[v[kk - 1]];
    zo[fi - 1] = dmin(dk, z[v[fi - 1]]);
    zo[fj - 1] = dmin(dk, z[v[fj - 1]]);
  }
  const int o = perm[i];
#pragma unroll
  for (int m = 0; m < 3; ++m) zout[3 * (size_t)o + m] = zo[m];
  if (pd) {
    pd[4 * (size_t)o] = part;
    pd[4 * (size_t)o + 1] = n;
    pd[4 * (size_t)o + 2] = fi;
    pd[4 * (size_t)o + 3] = fj;
  }
}

// fvsources.f90:27-32 for vertex p: the least-squares bed gradient over the
// vertices sharing an edge with p (lsq_gradient.f90:35-47, solver 1, with
// the weights 1/d^2 and normal-equation sums of lsq_setup.f90:32-80, type 2,
// as dataprepRDS.f90:93 sets them up), times -h.  zb: [2][P].
__global__ void k_node_grad(Geo g, const int* nn_ptr, const int* nn_idx, const double* z,
                            const double* h, double* zb) {
  const int p = xcd_tile() * blockDim.x + threadIdx.x;
  if (p >= g.P) return;
  const double4 q = g.vxyz[p];
  const int t0 = nn_ptr[p], t1 = nn_ptr[p + 1];
  double ai = ZERO, bi = ZERO, ci = ZERO, di = ZERO, ei = ZERO;
  for (int t = t0; t < t1; ++t) {
    const double4 l = g.vxyz[nn_idx[t]];
    const double dx = l.x - q.x, dy = l.y - q.y;
    const double w = ONE / (dx * dx + dy * dy);
    ai = ai + w * dx * dx;
    bi = bi + w * dx * dy;
    ci = ci + w * dy * dy;
  }
  const double det = ai * ci - bi * bi;
  const double zp = z[p];
  for (int t = t0; t < t1; ++t) {
    const int k = nn_idx[t];
    const double4 l = g.vxyz[k];
    const double dx = l.x - q.x, dy = l.y - q.y;
    const double w = ONE / (dx * dx + dy * dy);
    const double dphi = z[k] - zp;
    di = di + w * dphi * dx;
    ei = ei + w * dphi * dy;
  }
  const double zx = (ci * di - bi * ei) / det, zy = (ai * ei - di * bi) / det;
  zb[p] = -zx * h[p];
  zb[g.P + p] = -zy * h[p];
}

// fvsources.f90:34-44 for vertex p: source(:,2:3) = g times the median-dual
// integral of zb (integral_md.f90:33-56: per triangle of n2t, in its order,
// area * (c1 zb(p) + c2 (zb(j) + zb(k))), j and k the triangle's next
// vertices after p).  c1 and c2 are the routine's default-real literals;
// c1 = 2.037..., not the 11/54 of its comment.  Written in input vertex
// order, component-major: src[c*P + vperm[p]].
__global__ void k_node_integral(Geo g, const double* zb, const int* vperm, double* src) {
  const int p = xcd_tile() * blockDim.x + threadIdx.x;
  if (p >= g.P) return;
  const int N = g.N, P = g.P;
  const double c1 = (double)2.03703703703704f, c2 = (double)6.481481481481481E-2f;
  double ix = ZERO, iy = ZERO;
  for (int t = g.vptr[p]; t < g.vptr[p + 1]; ++t) {
    const int c = g.vcell[t];
    const int a0 = g.cvtx[c], a1 = g.cvtx[N + c], a2 = g.cvtx[2 * N + c];
    const int j = (a0 == p) ? a1 : (a1 == p) ? a2 : a0;
    const int k = (a0 == p) ? a2 : (a1 == p) ? a0 : a1;
    const double area = tri_area(load_tri(g, c));
    ix = ix + area * (c1 * zb[p] + c2 * (zb[j] + zb[k]));
    iy = iy + area * (c1 * zb[P + p] + c2 * (zb[P + j] + zb[P + k]));
  }
  const int o = vperm[p];
  src[o] = ZERO;
  src[P + o] = G * ix;
  src[2 * (size_t)P + o] = G * iy;
}

// ---------------------------------------------------------------- scratch
// The scratch buffer of storm_fvt_get_gradients_noalloc: delh of up to
// NOALLOC_MAX_CELLS cells (x then y components), 32 MiB.  Each device the
// library runs on holds its own copy until the process ends; every call on
// a device shares that device's copy.
constexpr int NOALLOC_MAX_CELLS = STORM_FVT_NOALLOC_MAX_CELLS;
static __device__ __align__(256) double scratch[2 * (size_t)NOALLOC_MAX_CELLS];

// ---------------------------------------------------------------- host side
#define HIPCHK(x)                                                                       \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) {                                                             \
      throw StormErr(STORM_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_));   \
    }                                                                                   \
  } while (0)

struct StormErr {
  int code;
  std::string msg;
  StormErr(int c, std::string m) : code(c), msg(std::move(m)) {}
};

template <class T>
T* dupload(std::vector<void*>& owned, const std::vector<T>& h) {
  T* d = nullptr;
  size_t n = std::max<size_t>(h.size(), 1);
  HIPCHK(hipMalloc(&d, n * sizeof(T)));
  owned.push_back(d);
  if (!h.empty()) HIPCHK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

template <class T>
T* dalloc(std::vector<void*>& owned, size_t n) {
  T* d = nullptr;
  HIPCHK(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T)));
  owned.push_back(d);
  HIPCHK(hipMemset(d, 0, std::max<size_t>(n, 1) * sizeof(T)));
  return d;
}

}  // namespace

namespace {
// The device address of `scratch` per device, taken once per device at
// context creation with that device current (the host symbol is only a
// stand-in for it).
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch[MAX_DEVICES] = {};
}  // namespace

struct storm_fvt {
  Prep prep;
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<void*> owned;
  Geo g{};
  Opts o{};
  State S{};
  double* slots[4] = {nullptr, nullptr, nullptr, nullptr};  // physical slot buffers
  double *d_h0 = nullptr, *d_u0 = nullptr, *d_v0 = nullptr;
  CellD *d_a = nullptr, *d_b = nullptr;  // stage state, double-buffered
  int* d_out_type = nullptr;
  int *d_qts_ptr = nullptr, *d_hts_ptr = nullptr;
  double *d_qts_t = nullptr, *d_qts_v = nullptr, *d_hts_t = nullptr, *d_hts_v = nullptr;
  // entries the Q(t) and H(t) table buffers hold (at least one per string,
  // so that storm_fvt_set_bc_values writes into them in place)
  size_t cap_q = 0, cap_h = 0;
  // the final to_vertices' tables and the post-loop fields (vertex
  // velocities, vertex drag coefficient, error estimates)
  OutTab tab{};
  double *d_uvtx = nullptr, *d_vvtx = nullptr, *d_cdvtx = nullptr, *d_eest = nullptr;
  // Where the vertex fields of get_vertex come from: dataprepFVT's (until a
  // step or set_state), the last finalize's, or none that is current (steps
  // or set_state since); vpass: hvtx on the device is the last stage's
  // (a step has run since the last set_state)
  enum { V_PREP, V_FINAL, V_NONE } vsrc = V_PREP;
  bool vpass = false;
  int* d_perm = nullptr;           // internal cell -> input cell
  // the node routines' tables, uploaded at their first call
  int *d_vperm = nullptr, *d_nn_ptr = nullptr, *d_nn_idx = nullptr;
  double* scratch_ptr = nullptr;   // this device's copy of `scratch`
  uint64_t* d_cmeta = nullptr;     // g.cmeta (culvert flags are set in it)
  // wind terms (storm_fvt_set_wind): per cell wcoef1, wcoef2
  double2* d_wind = nullptr;
  // culverts (storm_fvt_set_culverts): their input cells, device tables and
  // the buffers that hold them (replaced by the next call)
  std::vector<int> cv_in, cv_out;
  CvTab cv{};
  double* d_cvsrc = nullptr;
  std::vector<void*> cv_owned;
  // residual smoothing (opts rsavg): level schedule, factors, sweeps
  bool rsavg = false;
  RsTab rs{};
  double rs_fac1 = 0.0, rs_fac2 = 0.0;
  int rs_n = 0;
  // the shared-memory sweep (Prep::rs_fast): tables, the far values' sources,
  // phi0 in level order, the level-ordered results of alternate sweeps and
  // the old far values of the sweep being run
  bool rsf = false;
  int rsf_threads = 0;  // the widest level (and fill list) in whole waves
  RsFast rsf_tab{};
  const int *d_rsf_fsrc = nullptr, *d_rsf_fdst = nullptr;
  double4 *d_p0lo = nullptr, *d_lo_a = nullptr, *d_lo_b = nullptr, *d_fval = nullptr;
  double* d_rs_dummy = nullptr;
  Tiles tl{};
  int cap = 0;
  std::vector<Ctrl> hctrl;  // host copy of the reaches' control words
  std::vector<int> hvdone;  // zeros for State::vdone at the start of a run
  int iter_host = 0;
  bool started = false;
  bool finalized = false;
  double last_run_s = 0.0;
  // RK order 1 keeps two physical slots and swaps them every step
  int r1_flip = 0;
  // input state of the last stage run: to_vertices after the loop works on
  // the water surface formed at the start of that stage (solverFVT.f90:169,483)
  const CellD* last_din = nullptr;

  // two-step graph (opts.graph) and the buffer phase it was captured at
  hipGraphExec_t gexec = nullptr;
  const CellD* g_din = nullptr;
  int g_flip = 0;

  ~storm_fvt() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    for (void* p : owned) (void)hipFree(p);
    for (void* p : cv_owned) (void)hipFree(p);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  double* slot_ptr(int s) const {  // logical rk slot s -> buffer
    const int R = prep.R;
    if (R == 1) return slots[(s + r1_flip) & 1];
    return (s == 0) ? slots[R] : slots[s];
  }
};

namespace {

void host_time_bcs(const Prep& P, double time, double* qin, double* hbc) {
  for (int k = 0; k < P.n_out; ++k) {
    if (P.out_type[k] != 1) continue;
    int a = P.hts_ptr[k], n = P.hts_ptr[k + 1] - a;
    if (n <= 0) continue;
    const double* tl = &P.hts_t[a];
    const double* tv = &P.hts_v[a];
    if (time < tl[0]) hbc[k] = tv[0];
    else if (time >= tl[n - 1]) hbc[k] = tv[n - 1];
    else
      for (int i = 0; i < n - 1; ++i)
        if (time >= tl[i] && time < tl[i + 1]) {
          double d = time - tl[i], d2 = tl[i + 1] - tl[i];
          hbc[k] = tv[i] + d * (tv[i + 1] - tv[i]) / d2;
          break;
        }
  }
  for (int k = 0; k < P.n_in; ++k) {
    int a = P.qts_ptr[k], n = P.qts_ptr[k + 1] - a;
    if (n <= 0) continue;
    const double* tl = &P.qts_t[a];
    const double* tv = &P.qts_v[a];
    if (time < tl[0]) qin[k] = tv[0];
    else if (time >= tl[n - 1]) qin[k] = tv[n - 1];
    else
      for (int i = 0; i < n - 1; ++i)
        if (time >= tl[i] && time < tl[i + 1]) {
          double d = time - tl[i], d2 = tl[i + 1] - tl[i];
          qin[k] = tv[i] + d * (tv[i + 1] - tv[i]) / d2;
          break;
        }
  }
}

// The boundary tables on the device: buffers sized for the larger of the
// given tables and one entry per string, filled from the host copies.
void upload_tables(storm_fvt* C) {
  Prep& P = C->prep;
  auto& own = C->owned;
  C->cap_q = std::max<size_t>({P.qts_t.size(), (size_t)P.n_in, 1});
  C->cap_h = std::max<size_t>({P.hts_t.size(), (size_t)P.n_out, 1});
  C->d_out_type = dupload(own, P.out_type);
  C->d_qts_ptr = dupload(own, P.qts_ptr);
  C->d_hts_ptr = dupload(own, P.hts_ptr);
  C->d_qts_t = dalloc<double>(own, C->cap_q);
  C->d_qts_v = dalloc<double>(own, C->cap_q);
  C->d_hts_t = dalloc<double>(own, C->cap_h);
  C->d_hts_v = dalloc<double>(own, C->cap_h);
  auto put = [](double* d, const std::vector<double>& h) {
    if (!h.empty()) HIPCHK(hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
  };
  put(C->d_qts_t, P.qts_t); put(C->d_qts_v, P.qts_v);
  put(C->d_hts_t, P.hts_t); put(C->d_hts_v, P.hts_v);
}

void set_lds_limit(int cap, bool xsf) {
  const void* f = xsf ? reinterpret_cast<const void*>(&k_stage<256, true>)
                : (cap == 256) ? reinterpret_cast<const void*>(&k_stage<256>)
                : (cap == 384) ? reinterpret_cast<const void*>(&k_stage<384>)
                               : reinterpret_cast<const void*>(&k_stage<512>);
  HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)lds_bytes(cap, xsf)));
}

void setup_device(storm_fvt* C) {
  Prep& P = C->prep;
  auto& own = C->owned;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    throw StormErr(STORM_E_DEVICE, "no HIP device available");
  HIPCHK(hipGetDevice(&C->device));
  HIPCHK(hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&C->ev0));
  HIPCHK(hipEventCreate(&C->ev1));
  Geo& g = C->g;
  g.N = P.N; g.P = P.P; g.R = P.R;
  g.cvtx = dupload(own, P.cvtx); g.cnbr = dupload(own, P.cnbr);
  C->d_cmeta = dupload(own, P.cmeta);
  g.cmeta = C->d_cmeta;
  {
    std::vector<CellS> cs(P.N);
    for (int i = 0; i < P.N; ++i) cs[i] = CellS{P.xc[i], P.yc[i], P.z[i], P.cd[i]};
    g.cs = dupload(own, cs);
  }
  g.vxyz = reinterpret_cast<const double4*>(dupload(own, P.vxyz));
  g.mptr = dupload(own, P.mptr); g.mcell = dupload(own, P.mcell);
  g.gconn = dupload(own, P.gconn);
  g.gcx = dupload(own, P.gcx); g.gcy = dupload(own, P.gcy); g.gwt = dupload(own, P.gwt);
  g.vptr = dupload(own, P.vptr); g.vcell = dupload(own, P.vcell); g.vbc = dupload(own, P.vbc);
  g.vell = dupload(own, P.vell); g.vwt = dupload(own, P.vwt);
  g.vzb = reinterpret_cast<const double2*>(dupload(own, P.vzb));
  g.qs_ptr = dupload(own, P.qs_ptr); g.qs_cell = dupload(own, P.qs_cell);
  g.qs_slot = dupload(own, P.qs_slot);
  g.qs_rcx = dupload(own, P.qs_rcx); g.qs_rcy = dupload(own, P.qs_rcy);
  g.qs_nx = dupload(own, P.qs_nx); g.qs_ny = dupload(own, P.qs_ny);
  g.qs_len = dupload(own, P.qs_len);
  g.in_type = dupload(own, P.in_type);
  g.svptr = dupload(own, P.r_svptr); g.svtx = dupload(own, P.r_svtx);
  g.n_in = P.n_in; g.n_out = P.n_out; g.nr = P.nr;
  upload_tables(C);
  Opts& o = C->o;
  const storm_fvt_opts_t& q = P.o;
  o.rkorder = q.rkorder; o.dt_op = q.dt_op; o.steady = q.steady; o.stop_cr = q.stop_cr;
  o.btype = q.btype; o.hvalve = q.hvalve; o.limiter = q.limiter; o.flux = q.flux;
  o.gradient = q.gradient; o.friction = q.friction; o.vclip = q.vclip;
  o.nonfinite_stop = q.nonfinite_stop;
  o.xsf = q.fluxes == 1;
  o.wind = 0;
  o.ncv = 0;
  o.cfl = q.cfl; o.rstop = q.rstop; o.h_dry = q.h_dry; o.h_wet = q.h_wet; o.vceiling = q.vceiling;
  const size_t N = P.N;
  const int R = P.R;
  const int nslots = (R == 1) ? 2 : R + 1;
  for (int s = 0; s < nslots && s < 4; ++s)
    if (R == 1 || s >= 1) C->slots[s] = dalloc<double>(own, 3 * N);
  // tiles of the stage kernel
  C->cap = q.block > 0 ? q.block : DEFAULT_CAP;
  if (C->cap != 256 && C->cap != 384 && C->cap != 512)
    throw StormErr(STORM_E_ARG, "block (tile capacity) must be 256, 384 or 512");
  if (o.xsf && C->cap != 256)
    throw StormErr(STORM_E_ARG, "the xsf edge loop runs with tile capacity 256");
  if (q.rsavg && C->cap != 256)
    throw StormErr(STORM_E_ARG, "residual smoothing runs with tile capacity 256");
  if (build_tiles(P, C->cap) != STORM_OK) throw StormErr(STORM_E_MESH, P.err);
  Tiles& tl = C->tl;
  tl.hdr = reinterpret_cast<const int4*>(dupload(own, P.t_hdr));
  tl.hrec = reinterpret_cast<const int4*>(dupload(own, P.t_hrec));
  tl.hmeta = dupload(own, P.t_hmeta);
  tl.cln = dupload(own, P.t_cln);
  tl.ehdr = reinterpret_cast<const int2*>(dupload(own, P.t_ehdr));
  tl.erec = dupload(own, P.t_erec);
  tl.nt = P.nt;
  {
    // per 256-vertex chunk of the vertex pass, the tiles owning the cells
    // incident to its vertices (the vertex pass skips its gathers where all
    // of them are dry)
    std::vector<int> tile_of(N);
    for (int b = 0; b < P.nt; ++b)
      for (int i = P.t_ptr[b]; i < P.t_ptr[b + 1]; ++i) tile_of[i] = b;
    const int PP = P.P, nch = (PP + 255) / 256;
    std::vector<int> vc_ptr(nch + 1, 0), vc_tile, seen;
    for (int c = 0; c < nch; ++c) {
      seen.clear();
      for (int p = 256 * c; p < std::min(PP, 256 * (c + 1)); ++p)
        for (int t = 0; t < P.vW; ++t) {
          const int k = P.vell[(size_t)t * PP + p];
          if (k >= 0) seen.push_back(tile_of[k]);
        }
      std::sort(seen.begin(), seen.end());
      seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
      vc_tile.insert(vc_tile.end(), seen.begin(), seen.end());
      vc_ptr[c + 1] = (int)vc_tile.size();
    }
    if (vc_tile.empty()) vc_tile.push_back(0);
    C->g.vc_ptr = dupload(own, vc_ptr);
    C->g.vc_tile = dupload(own, vc_tile);
  }
  {
    std::vector<RInfo> ri(P.nr);
    for (int r = 0; r < P.nr; ++r) {
      RInfo& q = ri[r];
      q.in_base = P.r_in[r]; q.n_in = P.r_in[r + 1] - P.r_in[r];
      q.out_base = P.r_out[r]; q.n_out = P.r_out[r + 1] - P.r_out[r];
      q.tile0 = P.r_tile[r]; q.tile1 = P.r_tile[r + 1];
      q.last_cell = P.r_last[r];
      q.pad = 0;
      q.bch_max = P.r_bch[r];
      q.pad2 = 0.0;
    }
    C->g.reach = dupload(own, ri);
  }
  set_lds_limit(C->cap, C->o.xsf != 0);
  State& S = C->S;
  S.gwork = (q.gradient == STORM_GRAD_GAUSS) ? dalloc<double>(own, 6 * N) : nullptr;
  S.hvtx = dalloc<double>(own, P.P);
  S.zetavtx = dalloc<double>(own, P.P);
  S.vwet = dalloc<uint8_t>(own, P.P);
  S.phi = dalloc<double>(own, 3 * N);
  S.dry_post = dalloc<uint8_t>(own, N);
  S.zeta = dalloc<double>(own, N);
  S.twet = dalloc<uint8_t>(own, P.nt);
  S.part = dalloc<double>(own, PART * (size_t)P.nt);
  C->d_a = dalloc<CellD>(own, N);
  C->d_b = dalloc<CellD>(own, N);
  S.ctrl = dalloc<Ctrl>(own, P.nr);
  S.vdone = dalloc<int>(own, P.nr);
  S.sq = dalloc<double>(own, P.n_in);
  S.sh = dalloc<double>(own, P.n_out);
  C->d_h0 = dupload(own, P.h0); C->d_u0 = dupload(own, P.u0); C->d_v0 = dupload(own, P.v0);
  C->d_perm = dupload(own, P.perm);
  if (q.rsavg) {
    // solverFVT.f90:59-60 with epsRSAVG as read_options.f90:351 leaves it
    const double eps = std::min(std::fabs(q.rsavg_eps), 1.0);
    C->rsavg = true;
    C->rs_fac1 = 1.0 / (1.0 + eps * 3.0);
    C->rs_fac2 = eps / (1.0 + eps * 3.0);
    C->rs_n = q.rsavg_n;
    RsTab& rs = C->rs;
    rs.lptr = dupload(own, P.rs_lptr);
    rs.cell = dupload(own, P.rs_cell);
    rs.nptr = dupload(own, P.rs_nptr);
    rs.nbr = dupload(own, P.rs_nbr);
    rs.rlev = dupload(own, P.r_lev);
    S.phi0 = dalloc<double>(own, 3 * N);
    S.phis = (C->rs_n > 0) ? dalloc<double>(own, 3 * N) : S.phi0;
    if (P.rs_fast && C->rs_n > 0) {
      RsFast& T = C->rsf_tab;
      T.rlev = rs.rlev;
      {
        std::vector<int> cf(2 * N);
        for (size_t k = 0; k < N; ++k) { cf[2 * k] = P.rs_cell[k]; cf[2 * k + 1] = P.rs_farw[k]; }
        T.cf = reinterpret_cast<const int2*>(dupload(own, cf));
        std::vector<int16_t> code(RSF_E * N);
        for (size_t k = 0; k < N; ++k)
          for (int e = 0; e < RSF_E; ++e) code[RSF_E * k + e] = (int16_t)P.rs_code[e * N + k];
        T.code = reinterpret_cast<const int4*>(dupload(own, code));
        std::vector<int> pos(N);
        for (size_t k = 0; k < N; ++k) pos[P.rs_cell[k]] = (int)k;
        S.rs_pos = dupload(own, pos);
      }
      T.hdr = reinterpret_cast<const int4*>(dupload(own, P.rs_hdr));
      T.pptr = dupload(own, P.rs_pptr);
      T.psrc = dupload(own, P.rs_psrc);
      T.pdst = dupload(own, P.rs_pdst);
      T.N = (int)N;
      T.NF = (int)P.rs_fsrc.size();
      T.L = (int)P.rs_lptr.size() - 1;
      C->d_rsf_fsrc = dupload(own, P.rs_fsrc);
      C->d_rsf_fdst = dupload(own, P.rs_fdst);
      C->d_p0lo = reinterpret_cast<double4*>(dalloc<double>(own, 4 * N));
      C->d_lo_a = reinterpret_cast<double4*>(dalloc<double>(own, 4 * N));
      C->d_lo_b = reinterpret_cast<double4*>(dalloc<double>(own, 4 * N));
      C->d_fval = reinterpret_cast<double4*>(dalloc<double>(own, 4 * (size_t)std::max(T.NF, 1)));
      C->d_rs_dummy = dalloc<double>(own, 4);
      S.phis_lo = ((C->rs_n - 1) & 1) ? C->d_lo_b : C->d_lo_a;  // the last sweep's
      {
        // idle waves would still issue every load of a level, so the block
        // has only as many waves as the widest level or fill list needs
        int w = 1;
        for (size_t l = 0; l + 1 < P.rs_lptr.size(); ++l)
          w = std::max({w, P.rs_lptr[l + 1] - P.rs_lptr[l], P.rs_fptr[l + 1] - P.rs_fptr[l]});
        C->rsf_threads = std::min(RSF_NT, (w + 63) / 64 * 64);
      }
      HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rsavg_fast),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, RSF_LDS));
      C->rsf = true;
    }
  }
  {
    if (C->device < 0 || C->device >= MAX_DEVICES)
      throw StormErr(STORM_E_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[C->device]) {
      void* ptr = nullptr;
      HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
      g_scratch[C->device] = static_cast<double*>(ptr);
    }
    C->scratch_ptr = g_scratch[C->device];
  }
  C->err.reserve(256);  // error messages of the no-allocation entry point fit
  OutTab& T = C->tab;
  T.vo_ptr = dupload(own, P.vo_ptr);
  T.vwall = dupload(own, P.vwall);
  T.vo_nrm = reinterpret_cast<const double2*>(dupload(own, P.vo_nrm));
  T.wtang = reinterpret_cast<const double2*>(dupload(own, P.wtang));
  T.is_ptr = dupload(own, P.is_ptr);
  T.is_node = dupload(own, P.is_node);
  T.is_pn = reinterpret_cast<const double2*>(dupload(own, P.is_pn));
  T.is_len = dupload(own, P.is_len);
  C->d_uvtx = dalloc<double>(own, P.P);
  C->d_vvtx = dalloc<double>(own, P.P);
  C->d_cdvtx = dalloc<double>(own, P.P);
  C->d_eest = dalloc<double>(own, N);
}

template <int CAP>
void launch_init(storm_fvt* C) {
  hipLaunchKernelGGL(k_init<CAP>, dim3(C->tl.nt), dim3(CAP), 0, C->stream, C->g, C->o, C->tl,
                     C->slot_ptr(0), C->d_h0, C->d_u0, C->d_v0, C->d_a, C->S.zeta, C->S.twet,
                     C->S.part);
}

template <int CAP, bool XSF = false, int MODE = 0>
void launch_stage(storm_fvt* C, const Stage& st) {
  hipLaunchKernelGGL((k_stage<CAP, XSF, MODE>), dim3(C->tl.nt), dim3(CAP),
                     MODE == 2 ? 0 : lds_bytes(CAP, XSF), C->stream, C->g, C->o, C->tl, st, C->S);
}

// The stage with residual smoothing: the residual of every cell, the sweeps,
// then the update from the smoothed residual.
void launch_smoothed_stage(storm_fvt* C, const Stage& st) {
  if (C->o.xsf) launch_stage<256, true, 1>(C, st);
  else launch_stage<256, false, 1>(C, st);
  if (C->rs_n > 0 && C->rsf) {
    const RsFast& T = C->rsf_tab;
    hipLaunchKernelGGL(k_rs_order, dim3((T.N + 255) / 256), dim3(256), 0, C->stream, T,
                       (const double*)C->S.phi0, C->d_p0lo);
    const double4* src = C->d_p0lo;
    for (int sw = 0; sw < C->rs_n; ++sw) {
      double4* out_lo = (sw & 1) ? C->d_lo_b : C->d_lo_a;
      if (T.NF > 0)
        hipLaunchKernelGGL(k_rs_far, dim3((T.NF + 255) / 256), dim3(256), 0, C->stream, T,
                           C->d_rsf_fsrc, C->d_rsf_fdst, src, C->d_fval);
      hipLaunchKernelGGL(k_rsavg_fast, dim3(C->prep.nr), dim3(C->rsf_threads), RSF_LDS, C->stream, T,
                         (const Ctrl*)C->S.ctrl, C->rs_fac1, C->rs_fac2,
                         (const double4*)C->d_p0lo, src, (const double4*)C->d_fval, out_lo,
                         C->d_rs_dummy);
      src = out_lo;
    }
  } else if (C->rs_n > 0) {
    hipLaunchKernelGGL(k_rsavg, dim3(C->prep.nr), dim3(RS_THREADS), 0, C->stream, C->rs,
                       C->S.ctrl, C->prep.N, C->rs_n, C->rs_fac1, C->rs_fac2, C->S.phi0,
                       C->S.phis);
  }
  launch_stage<256, false, 2>(C, st);
}

void start_run(storm_fvt* C) {
  // solverFVT.f90:53-127
  Prep& P = C->prep;
  C->hctrl.assign(P.nr, Ctrl{});
  std::vector<double> qin(std::max(P.n_in, 1), 0.0), hbc(std::max(P.n_out, 1), 0.0);
  host_time_bcs(P, 0.0, qin.data(), hbc.data());
  HIPCHK(hipMemcpyAsync(C->S.ctrl, C->hctrl.data(), sizeof(Ctrl) * P.nr, hipMemcpyHostToDevice,
                        C->stream));
  C->hvdone.assign(P.nr, 0);
  HIPCHK(hipMemcpyAsync(C->S.vdone, C->hvdone.data(), sizeof(int) * P.nr, hipMemcpyHostToDevice,
                        C->stream));
  if (P.n_in)
    HIPCHK(hipMemcpyAsync(C->S.sq, qin.data(), sizeof(double) * P.n_in, hipMemcpyHostToDevice,
                          C->stream));
  if (P.n_out)
    HIPCHK(hipMemcpyAsync(C->S.sh, hbc.data(), sizeof(double) * P.n_out, hipMemcpyHostToDevice,
                          C->stream));
  C->r1_flip = 0;
  switch (C->cap) {
    case 256: launch_init<256>(C); break;
    case 384: launch_init<384>(C); break;
    default: launch_init<512>(C); break;
  }
  HIPCHK(hipGetLastError());
  C->S.din = C->d_a;
  C->S.dout = C->d_b;
  C->last_din = C->d_a;
  C->iter_host = 0;
  C->started = true;
  C->finalized = false;
}

// the SSP-RK Shu-Osher tables of dataprepFVT.f90:256-273
void rk_table(int R, double a[3][3], double b[3]) {
  std::memset(a, 0, sizeof(double) * 9);
  std::memset(b, 0, sizeof(double) * 3);
  if (R == 1) { a[0][0] = 1.0; b[0] = 1.0; }
  else if (R == 2) { a[0][0] = 1.0; b[0] = 1.0; a[1][0] = 0.5; a[1][1] = 0.5; b[1] = 0.5; }
  else {
    a[0][0] = 1.0;
    a[1][0] = 0.750; a[1][1] = 0.250;
    a[2][0] = 1.0 / 3.0; a[2][1] = 0.0; a[2][2] = 2.0 / 3.0;
    b[0] = 1.0; b[1] = 0.250; b[2] = 2.0 / 3.0;
  }
}

CtrlArgs ctrl_args(storm_fvt* C, int begin) {
  CtrlArgs A;
  A.ctrl = C->S.ctrl; A.part = C->S.part; A.sq = C->S.sq; A.sh = C->S.sh; A.hvtx = C->S.hvtx;
  A.begin = begin;
  A.out_type = C->d_out_type; A.qts_ptr = C->d_qts_ptr; A.hts_ptr = C->d_hts_ptr;
  A.qts_t = C->d_qts_t; A.qts_v = C->d_qts_v; A.hts_t = C->d_hts_t; A.hts_v = C->d_hts_v;
  return A;
}

void launch_control(storm_fvt* C, int begin) {
  hipLaunchKernelGGL(k_control, dim3(C->prep.nr), dim3(CTRL_THREADS), 0, C->stream, C->g, C->o,
                     ctrl_args(C, begin));
}

// One time step: per stage the vertex pass and the stage kernel, where the
// first stage's vertex pass launch also runs the step's k_control (both
// need only the previous step's last stage).  With launch == false only the
// host's view of the buffers advances (after a graph replay of the same
// step).
void launch_step(storm_fvt* C, bool launch = true) {
  Prep& P = C->prep;
  const int R = P.R;
  const int N = P.N;
  double a[3][3], b[3];
  rk_table(R, a, b);
  for (int s = 1; s <= R; ++s) {
    Stage st{};
    for (int j = 0; j < s; ++j) { st.slot[j] = C->slot_ptr(j); st.alpha[j] = a[s - 1][j]; }
    st.out = C->slot_ptr(s);
    st.hin = C->slot_ptr(s - 1);
    st.beta = b[s - 1];
    st.s = s;
    st.last = (s == R);
    if (launch) {
      const int nvb = (P.P + 255) / 256;
      if (s == 1)
        hipLaunchKernelGGL(k_vertex_first, dim3(P.nr + nvb), dim3(256), 0, C->stream, C->g, C->o,
                           C->S.zeta, C->S.twet, C->S.vwet, C->S.vdone, ctrl_args(C, 1));
      else
        hipLaunchKernelGGL(k_vertex_stage, dim3(nvb), dim3(256), 0, C->stream, C->g, C->o,
                           C->S.zeta, C->S.twet, C->S.hvtx, C->S.vwet, C->S.ctrl, C->S.sh);
      if (C->o.gradient == STORM_GRAD_GAUSS && C->o.limiter != 0)
        hipLaunchKernelGGL(k_gauss1, dim3((N + 255) / 256), dim3(256), 0, C->stream, C->g,
                           C->S.din, C->S.gwork);
      if (C->o.ncv)
        hipLaunchKernelGGL(k_cv_src, dim3((C->o.ncv + 63) / 64), dim3(64), 0, C->stream, C->g,
                           C->o, C->S.din, C->S.gwork, C->cv, C->d_cvsrc);
      if (C->rsavg) launch_smoothed_stage(C, st);
      else if (C->o.xsf) launch_stage<256, true>(C, st);
      else switch (C->cap) {
        case 256: launch_stage<256>(C, st); break;
        case 384: launch_stage<384>(C, st); break;
        default: launch_stage<512>(C, st); break;
      }
    }
    C->last_din = C->S.din;
    CellD* next = C->S.dout;
    C->S.dout = const_cast<CellD*>(C->S.din);
    C->S.din = next;
  }
  if (R == 1) C->r1_flip ^= 1;
}

// Two time steps captured as one HIP graph.  Two steps bring the buffer
// roles (stage state double buffer, RK1 slot flip) back to where they were,
// so the graph replays from any even step after the captured one.  No
// memset node is recorded: the step launches kernels only.
void ensure_graph(storm_fvt* C) {
  if (C->gexec) return;
  const CellD* din = C->S.din;
  CellD* dout = C->S.dout;
  const CellD* last = C->last_din;
  const int flip = C->r1_flip;
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(C->stream, hipStreamCaptureModeThreadLocal));
  launch_step(C);
  launch_step(C);
  const hipError_t e = hipStreamEndCapture(C->stream, &graph);
  // the capture ran nothing: restore the host's view of the buffers
  C->S.din = din; C->S.dout = dout; C->last_din = last; C->r1_flip = flip;
  HIPCHK(e);
  const hipError_t ei = hipGraphInstantiate(&C->gexec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIPCHK(ei);
  C->g_din = din;
  C->g_flip = flip;
}

int set_err(storm_fvt* C, int code, const std::string& msg) {
  if (C) C->err = msg;
  return code;
}
// A literal message: assigned into the context's reserved string, so this
// error path allocates nothing.
int set_err(storm_fvt* C, int code, const char* msg) {
  if (C) C->err.assign(msg);
  return code;
}

// Device buffers of one call, freed on every way out.
struct CallBufs {
  std::vector<void*> p;
  ~CallBufs() { for (void* q : p) (void)hipFree(q); }
  template <class T>
  T* get(size_t n) {
    T* d = nullptr;
    HIPCHK(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T)));
    p.push_back(d);
    return d;
  }
};

// Vertex values in input order -> a device array in internal order.
double* vertex_upload(storm_fvt* C, CallBufs& B, const double* a) {
  const Prep& P = C->prep;
  std::vector<double> t(P.P);
  for (int pp = 0; pp < P.P; ++pp) t[pp] = a[P.vperm[pp]];
  double* d = B.get<double>(P.P);
  HIPCHK(hipMemcpy(d, t.data(), sizeof(double) * P.P, hipMemcpyHostToDevice));
  return d;
}

}  // namespace

// ---------------------------------------------------------------- C ABI
extern "C" {

void storm_fvt_default_opts(storm_fvt_opts_t* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->rkorder = 2; o->dt_op = 0; o->steady = 1; o->stop_cr = 0; o->istop = 0; o->btype = 0;
  o->hvalve = 1; o->limiter = 2; o->flux = STORM_FLUX_AC; o->gradient = STORM_GRAD_LSQ;
  o->friction = 0; o->vclip = 0;
  o->cfl = 1.0; o->rstop = (double)1.0e-4f; o->h_dry = 0.001; o->h_wet = 0.001; o->vceiling = 35.0;
  o->reorder = 1; o->block = 0; o->graph = 0;
  o->rsavg = 0; o->rsavg_n = 2; o->rsavg_eps = 0.5;
}

static int create_ctx(const storm_mesh_t* mesh, const storm_fvt_opts_t* opts,
                      const ReachSplit* split, storm_fvt_t** ctx) {
  if (!ctx) return STORM_E_ARG;
  *ctx = nullptr;
  std::unique_ptr<storm_fvt> C(new (std::nothrow) storm_fvt());
  if (!C) return STORM_E_NOMEM;
  try {
    int rc = build_prep(mesh, opts, C->prep, split);
    if (rc != STORM_OK) {
      std::fprintf(stderr, "storm_fvt_create: %s\n", C->prep.err.c_str());
      return rc;
    }
    setup_device(C.get());
  } catch (const StormErr& e) {
    std::fprintf(stderr, "storm_fvt_create: %s\n", e.msg.c_str());
    return e.code;
  } catch (const std::bad_alloc&) {
    return STORM_E_NOMEM;
  }
  *ctx = C.release();
  return STORM_OK;
}

int storm_fvt_create(const storm_mesh_t* mesh, const storm_fvt_opts_t* opts, storm_fvt_t** ctx) {
  return create_ctx(mesh, opts, nullptr, ctx);
}

int storm_fvt_create_batch(int32_t n, const storm_mesh_t* meshes, const storm_fvt_opts_t* opts,
                           storm_fvt_t** ctx) {
  if (!ctx) return STORM_E_ARG;
  *ctx = nullptr;
  if (n < 1 || !meshes || !opts) return STORM_E_ARG;
  // A reach that stops runs no stage while the others' steps go on swapping
  // the buffer roles.  Under an odd RK order a step swaps them an odd number
  // of times (RK1 also flips its slots), so the stopped reach's state would
  // end in the buffer the others no longer read as the state: batched steady
  // runs take an even order.
  if (n > 1 && opts->steady && (opts->rkorder & 1)) return STORM_E_ARG;
  // the reaches laid end to end as one mesh: ids offset, strings and tables
  // appended in reach order
  const storm_mesh_t& m0 = meshes[0];
  int64_t np = 0, ne = 0, ned = 0, nin = 0, nout = 0, nin_nodes = 0, nout_nodes = 0, nqt = 0, nht = 0;
  bool edges = true;
  for (int r = 0; r < n; ++r) {
    const storm_mesh_t& m = meshes[r];
    if (m.n_pts < 3 || m.n_elems < 1 || m.n_in < 0 || m.n_out < 0 || !m.x || !m.y || !m.tri ||
        !m.z || !m.h || !m.u || !m.v || !m.cd || m.z_at_vtx != m0.z_at_vtx ||
        m.h_at_vtx != m0.h_at_vtx || m.u_at_vtx != m0.u_at_vtx || m.v_at_vtx != m0.v_at_vtx ||
        m.cd_at_vtx != m0.cd_at_vtx)
      return STORM_E_ARG;
    np += m.n_pts; ne += m.n_elems;
    edges = edges && m.n_edges > 0 && m.edge_p && m.edge_e;
    ned += m.n_edges;
    nin += m.n_in; nout += m.n_out;
    if (m.n_in) { nin_nodes += m.in_ptr[m.n_in]; nqt += m.qts_ptr[m.n_in]; }
    if (m.n_out) { nout_nodes += m.out_ptr[m.n_out]; nht += m.hts_ptr[m.n_out]; }
  }
  if (np > INT32_MAX / 2 || ne > INT32_MAX / 4) return STORM_E_ARG;
  std::vector<double> x, y, z, h, u, v, cd, qt, qv, ht, hv;
  std::vector<int32_t> tri, ep, ee, in_type, in_ptr{0}, in_nodes, qts_ptr{0}, out_type, out_ptr{0},
      out_nodes, hts_ptr{0};
  x.reserve(np); y.reserve(np); tri.reserve(3 * ne);
  ReachSplit sp;
  sp.cell_ptr = {0}; sp.in_ptr = {0}; sp.out_ptr = {0};
  int32_t po = 0, co = 0;
  auto app = [](std::vector<double>& d, const double* s, int k) { d.insert(d.end(), s, s + k); };
  for (int r = 0; r < n; ++r) {
    const storm_mesh_t& m = meshes[r];
    app(x, m.x, m.n_pts); app(y, m.y, m.n_pts);
    for (int t = 0; t < 3 * m.n_elems; ++t) tri.push_back(m.tri[t] + po);
    if (edges)
      for (int t = 0; t < 2 * m.n_edges; ++t) {
        ep.push_back(m.edge_p[t] + po);
        ee.push_back(m.edge_e[t] < 0 ? -1 : m.edge_e[t] + co);
      }
    app(z, m.z, m.z_at_vtx ? m.n_pts : m.n_elems);
    app(h, m.h, m.h_at_vtx ? m.n_pts : m.n_elems);
    app(u, m.u, m.u_at_vtx ? m.n_pts : m.n_elems);
    app(v, m.v, m.v_at_vtx ? m.n_pts : m.n_elems);
    app(cd, m.cd, m.cd_at_vtx ? m.n_pts : m.n_elems);
    for (int k = 0; k < m.n_in; ++k) {
      in_type.push_back(m.in_type[k]);
      for (int t = m.in_ptr[k]; t < m.in_ptr[k + 1]; ++t) in_nodes.push_back(m.in_nodes[t] + po);
      in_ptr.push_back((int32_t)in_nodes.size());
      for (int t = m.qts_ptr[k]; t < m.qts_ptr[k + 1]; ++t) {
        qt.push_back(m.qts_t[t]);
        qv.push_back(m.qts_v[t]);
      }
      qts_ptr.push_back((int32_t)qt.size());
    }
    for (int k = 0; k < m.n_out; ++k) {
      out_type.push_back(m.out_type[k]);
      for (int t = m.out_ptr[k]; t < m.out_ptr[k + 1]; ++t) out_nodes.push_back(m.out_nodes[t] + po);
      out_ptr.push_back((int32_t)out_nodes.size());
      for (int t = m.hts_ptr[k]; t < m.hts_ptr[k + 1]; ++t) {
        ht.push_back(m.hts_t[t]);
        hv.push_back(m.hts_v[t]);
      }
      hts_ptr.push_back((int32_t)ht.size());
    }
    po += m.n_pts; co += m.n_elems;
    sp.cell_ptr.push_back(co);
    sp.in_ptr.push_back((int)in_type.size());
    sp.out_ptr.push_back((int)out_type.size());
  }
  (void)nin; (void)nout; (void)nin_nodes; (void)nout_nodes; (void)nqt; (void)nht; (void)ned;
  storm_mesh_t M{};
  M.n_pts = po; M.n_elems = co;
  M.x = x.data(); M.y = y.data(); M.tri = tri.data();
  M.n_edges = edges ? (int32_t)(ep.size() / 2) : 0;
  M.edge_p = edges ? ep.data() : nullptr;
  M.edge_e = edges ? ee.data() : nullptr;
  M.z_at_vtx = m0.z_at_vtx; M.h_at_vtx = m0.h_at_vtx; M.u_at_vtx = m0.u_at_vtx;
  M.v_at_vtx = m0.v_at_vtx; M.cd_at_vtx = m0.cd_at_vtx;
  M.z = z.data(); M.h = h.data(); M.u = u.data(); M.v = v.data(); M.cd = cd.data();
  M.n_in = (int32_t)in_type.size();
  M.in_type = in_type.data(); M.in_ptr = in_ptr.data(); M.in_nodes = in_nodes.data();
  M.qts_ptr = qts_ptr.data(); M.qts_t = qt.data(); M.qts_v = qv.data();
  M.n_out = (int32_t)out_type.size();
  M.out_type = out_type.data(); M.out_ptr = out_ptr.data(); M.out_nodes = out_nodes.data();
  M.hts_ptr = hts_ptr.data(); M.hts_t = ht.data(); M.hts_v = hv.data();
  return create_ctx(&M, opts, &sp, ctx);
}

int storm_fvt_reach_status(storm_fvt_t* C, int32_t* iter, double* ttime, double* rmax,
                           uint8_t* done) {
  if (!C) return STORM_E_ARG;
  try {
    if (!C->started) start_run(C);
    const int nr = C->prep.nr;
    HIPCHK(hipMemcpyAsync(C->hctrl.data(), C->S.ctrl, sizeof(Ctrl) * nr, hipMemcpyDeviceToHost,
                          C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    for (int r = 0; r < nr; ++r) {
      const Ctrl& c = C->hctrl[r];
      if (iter) iter[r] = c.iter;
      if (ttime) ttime[r] = c.ttime;
      if (rmax) rmax[r] = c.rmax;
      if (done) done[r] = (uint8_t)(c.done != 0);
    }
  } catch (const StormErr& e) {
    return set_err(C, e.code, e.msg);
  }
  return STORM_OK;
}

int storm_fvt_set_state(storm_fvt_t* C, const double* h, const double* u, const double* v) {
  if (!C || !h || !u || !v) return set_err(C, STORM_E_ARG, "null argument");
  try {
    Prep& P = C->prep;
    for (int ii = 0; ii < P.N; ++ii) {
      int i = P.perm[ii];
      P.h0[ii] = h[i]; P.u0[ii] = u[i]; P.v0[ii] = v[i];
    }
    HIPCHK(hipMemcpyAsync(C->d_h0, P.h0.data(), sizeof(double) * P.N, hipMemcpyHostToDevice, C->stream));
    HIPCHK(hipMemcpyAsync(C->d_u0, P.u0.data(), sizeof(double) * P.N, hipMemcpyHostToDevice, C->stream));
    HIPCHK(hipMemcpyAsync(C->d_v0, P.v0.data(), sizeof(double) * P.N, hipMemcpyHostToDevice, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    C->started = false;
    C->vsrc = storm_fvt::V_NONE;
    C->vpass = false;
  } catch (const StormErr& e) {
    return set_err(C, e.code, e.msg);
  }
  return STORM_OK;
}

int storm_fvt_set_bc_values(storm_fvt_t* C, const double* qin, const double* hbc) {
  if (!C) return STORM_E_ARG;
  try {
    Prep& P = C->prep;
    // no launched step may still read the tables or the current values
    HIPCHK(hipStreamSynchronize(C->stream));
    if (qin) {
      P.qts_ptr.assign(P.n_in + 1, 0);
      for (int k = 0; k <= P.n_in; ++k) P.qts_ptr[k] = k;
      P.qts_t.assign(P.n_in, 0.0);
      P.qts_v.assign(qin, qin + P.n_in);
    }
    if (hbc) {
      P.hts_ptr.assign(P.n_out + 1, 0);
      for (int k = 0; k <= P.n_out; ++k) P.hts_ptr[k] = k;
      P.hts_t.assign(P.n_out, 0.0);
      P.hts_v.assign(hbc, hbc + P.n_out);
    }
    // in place: the buffers hold one entry per string (upload_tables), and a
    // captured graph keeps their addresses
    auto put = [](void* d, const void* h, size_t bytes) {
      if (bytes) HIPCHK(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
    };
    if (qin) {
      put(C->d_qts_ptr, P.qts_ptr.data(), sizeof(int) * P.qts_ptr.size());
      put(C->d_qts_t, P.qts_t.data(), sizeof(double) * P.n_in);
      put(C->d_qts_v, P.qts_v.data(), sizeof(double) * P.n_in);
    }
    if (hbc) {
      put(C->d_hts_ptr, P.hts_ptr.data(), sizeof(int) * P.hts_ptr.size());
      put(C->d_hts_t, P.hts_t.data(), sizeof(double) * P.n_out);
      put(C->d_hts_v, P.hts_v.data(), sizeof(double) * P.n_out);
    }
    // A run in progress goes on from its current step with the new values:
    // they replace the current ones now (get_time_bcs keeps them from the
    // next step's end on, and a steady run never calls it again)
    if (C->started) {
      if (qin) put(C->S.sq, qin, sizeof(double) * P.n_in);
      if (hbc && P.n_out) {
        std::vector<double> sh(P.n_out);
        HIPCHK(hipMemcpy(sh.data(), C->S.sh, sizeof(double) * P.n_out, hipMemcpyDeviceToHost));
        for (int k = 0; k < P.n_out; ++k)
          if (P.out_type[k] == 1) sh[k] = hbc[k];
        put(C->S.sh, sh.data(), sizeof(double) * P.n_out);
      }
    }
  } catch (const StormErr& e) {
    return set_err(C, e.code, e.msg);
  }
  return STORM_OK;
}

int storm_fvt_run(storm_fvt_t* C, int32_t nsteps, int32_t* iter, double* ttime, double* rmax) {
  if (!C || nsteps < 0) return set_err(C, STORM_E_ARG, "bad argument");
  try {
    if (!C->started) start_run(C);
    if (C->finalized) return set_err(C, STORM_E_STATE, "run after finalize: call set_state first");
    const int room = std::max(0, C->prep.o.istop - C->iter_host);
    const int n = std::min(nsteps, room);
    // The host's view of the buffer roles at the start of the call.  Steps
    // launched after a reach stopped (steady convergence, a non-finite stop)
    // run no stage, so they write no buffer: the view is rebuilt below from
    // the steps the reach took.
    const CellD* din0 = C->S.din;
    CellD* dout0 = C->S.dout;
    const CellD* last0 = C->last_din;
    const int flip0 = C->r1_flip;
    const int iter0 = C->hctrl[0].iter;
    HIPCHK(hipEventRecord(C->ev0, C->stream));
    int k = 0;
    if (C->prep.o.graph && n >= 2) {
      ensure_graph(C);
      if (C->S.din != C->g_din || C->r1_flip != C->g_flip) {  // one step to the graph's phase
        launch_step(C);
        ++k;
      }
      for (; k + 2 <= n; k += 2) {
        HIPCHK(hipGraphLaunch(C->gexec, C->stream));
        launch_step(C, false);
        launch_step(C, false);
      }
    }
    for (; k < n; ++k) launch_step(C);
    if (n > 0) launch_control(C, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(C->ev1, C->stream));
    const int nr = C->prep.nr;
    HIPCHK(hipMemcpyAsync(C->hctrl.data(), C->S.ctrl, sizeof(Ctrl) * nr, hipMemcpyDeviceToHost,
                          C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, C->ev0, C->ev1));
    C->last_run_s = ms * 1e-3;
    // steps taken so far: the most any reach has taken; once every reach
    // has converged no further step is launched
    bool all_done = true;
    C->iter_host = 0;
    for (const Ctrl& hc : C->hctrl) {
      C->iter_host = std::max(C->iter_host, hc.iter);
      all_done = all_done && hc.done;
    }
    if (all_done) C->iter_host = C->prep.o.istop;
    bool moved = false;  // some reach took a step in this call
    for (int r = 0; r < nr; ++r) moved = moved || C->hctrl[r].iter != 0;
    if (moved && n > 0) { C->vsrc = storm_fvt::V_NONE; C->vpass = true; }
    if (C->prep.nr == 1 && C->hctrl[0].iter - iter0 < n) {
      C->S.din = din0;
      C->S.dout = dout0;
      C->last_din = last0;
      C->r1_flip = flip0;
      for (int j = iter0; j < C->hctrl[0].iter; ++j) launch_step(C, false);
    }
    const Ctrl& h0 = C->hctrl[0];
    if (iter) *iter = h0.iter;
    if (ttime) *ttime = h0.ttime;
    if (rmax) *rmax = h0.rmax;
  } catch (const StormErr& e) {
    return set_err(C, e.code, e.msg);
  }
  return STORM_OK;
}

int storm_fvt_finalize(storm_fvt_t* C) {
  if (!C) return STORM_E_ARG;
  try {
    if (!C->started) start_run(C);
    Prep& P = C->prep;
    const int bs = 256;
    const dim3 gv((P.P + bs - 1) / bs), gc((P.N + bs - 1) / bs);
    // to_vertices after the loop (solverFVT.f90:483): the water surface of
    // the last stage's start and the dry flags of the post-step
    // find_dry_cells (:335); then ctr2vtx of cd (:484) and error_est(2)
    // (:487) on the last stage's delh
    const bool stepped = C->iter_host > 0;
    const uint8_t* dry = stepped ? C->S.dry_post : (const uint8_t*)nullptr;
    CellD* dfin = const_cast<CellD*>(C->S.din);
    hipLaunchKernelGGL(k_vertex, gv, dim3(bs), 0, C->stream, C->g, C->o, C->last_din, dry,
                       C->S.hvtx, C->S.vwet, C->S.zetavtx, (const double*)nullptr);
    hipLaunchKernelGGL(k_final_cells, gc, dim3(bs), 0, C->stream, C->g, dry, C->S.vwet, dfin);
    hipLaunchKernelGGL(k_out_vertex, gv, dim3(bs), 0, C->stream, C->g, C->o, dfin, dry, C->tab,
                       C->S.sh, C->S.hvtx, C->d_uvtx, C->d_vvtx, C->d_cdvtx);
    hipLaunchKernelGGL(k_out_inflow, dim3(P.nr), dim3(bs), 0, C->stream, C->g, C->o, C->tab,
                       C->S.sq, C->S.hvtx, C->d_uvtx, C->d_vvtx);
    if (stepped && C->o.gradient == STORM_GRAD_GAUSS && C->o.limiter != 0)
      hipLaunchKernelGGL(k_gauss1, gc, dim3(bs), 0, C->stream, C->g, C->last_din, C->S.gwork);
    hipLaunchKernelGGL(k_out_cells, gc, dim3(bs), 0, C->stream, C->g, C->o, C->last_din,
                       C->S.gwork, (const CellD*)dfin, (int)stepped, C->d_eest);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(C->stream));
    C->finalized = true;
    C->vsrc = storm_fvt::V_FINAL;
  } catch (const StormErr& e) {
    return set_err(C, e.code, e.msg);
  }
  return STORM_OK;
}

static int copy_cells(storm_fvt_t* C, const double* d, double* out) {
  Prep& P = C->prep;
  std::vector<double> tmp(P.N);
  HIPCHK(hipMemcpy(tmp.data(), d, sizeof(double) * P.N, hipMemcpyDeviceToHost));
  for (int ii = 0; ii < P.N; ++ii) out[P.perm[ii]] = tmp[ii];
  return STORM_OK;
}

int storm_fvt_get_state(storm_fvt_t* C, double* h, double* u, double* v, double* phi,
                        uint8_t* drycell) {
  if (!C) return STORM_E_ARG;
  try {
    if (!C->started) start_run(C);
    HIPCHK(hipStreamSynchronize(C->stream));
    Prep& P = C->prep;
    const size_t N = P.N;
    if (h) copy_cells(C, C->slot_ptr(0), h);
    std::vector<CellD> d;
    if (u || v || (drycell && C->iter_host == 0)) {
      d.resize(N);
      HIPCHK(hipMemcpy(d.data(), C->S.din, sizeof(CellD) * N, hipMemcpyDeviceToHost));
    }
    if (u)
      for (size_t ii = 0; ii < N; ++ii) u[P.perm[ii]] = d[ii].u;
    if (v)
      for (size_t ii = 0; ii < N; ++ii) v[P.perm[ii]] = d[ii].v;
    if (phi)
      for (int c = 0; c < 3; ++c) copy_cells(C, C->S.phi + c * N, phi + c * N);
    if (drycell) {
      // the flags of the post-step find_dry_cells (solverFVT.f90:335)
      if (C->iter_host > 0) {
        std::vector<uint8_t> tmp(N);
        HIPCHK(hipMemcpy(tmp.data(), C->S.dry_post, N, hipMemcpyDeviceToHost));
        for (size_t ii = 0; ii < N; ++ii) drycell[P.perm[ii]] = tmp[ii];
      } else {
        for (size_t ii = 0; ii < N; ++ii) drycell[P.perm[ii]] = d[ii].dry != 0.0;
      }
    }
  } catch (const StormErr& e) {
    return set_err(C, e.code, e.msg);
  }
  return STORM_OK;
}

int storm_fvt_get_conserved(storm_fvt_t* C, double* qh, double* qu, double* qv) {
  if (!C) return STORM_E_ARG;
  try {
    if (!C->started) start_run(C);
    HIPCHK(hipStreamSynchronize(C->stream));
    Prep& P = C->prep;
    const double* s = C->slot_ptr(0);
    if (qh) copy_cells(C, s, qh);
    if (qu) copy_cells(C, s + P.N, qu);
    if (qv) copy_cells(C, s + 2 * (size_t)P.N, qv);
  } catch (const StormErr& e) {
    return set_err(C, e.code, e.msg);
  }
  return STORM_OK;
}

static int copy_vertices(storm_fvt_t* C, const double* d, double* out) {
  Prep& P = C->prep;
  std::vector<double> tmp(P.P);
  HIPCHK(hipMemcpy(tmp.data(), d, sizeof(double) * P.P, hipMemcpyDeviceToHost));
  for (int pp = 0; pp < P.P; ++pp) out[P.vperm[pp]] = tmp[pp];
  return STORM_OK;
}

int storm_fvt_get_vertex(storm_fvt_t* C, double* hvtx, double* uvtx, double* vvtx,
                         double* zetavtx) {
  if (!C) return STORM_E_ARG;
  try {
    Prep& P = C->prep;
    if (C->vsrc == storm_fvt::V_PREP) {  // what dataprepFVT formed
      dataprep_fields(P, nullptr, nullptr, nullptr, nullptr, nullptr, hvtx, uvtx, vvtx, zetavtx);
      return STORM_OK;
    }
    if (C->vsrc == storm_fvt::V_NONE && ((uvtx || vvtx || zetavtx) || (hvtx && !C->vpass)))
      return set_err(C, STORM_E_STATE,
                     "the vertex fields are formed by storm_fvt_finalize (steps or set_state since)");
    HIPCHK(hipStreamSynchronize(C->stream));
    if (hvtx) copy_vertices(C, C->S.hvtx, hvtx);
    if (uvtx) copy_vertices(C, C->d_uvtx, uvtx);
    if (vvtx) copy_vertices(C, C->d_vvtx, vvtx);
    if (zetavtx) copy_vertices(C, C->S.zetavtx, zetavtx);
  } catch (const StormErr& e) {
    return set_err(C, e.code, e.msg);
  }
  return STORM_OK;
}

int storm_fvt_get_prep_fields(storm_fvt_t* C, double* z, double* zvtx, double* zeta, double* cd,
                              double* cdvtx) {
  if (!C) return STORM_E_ARG;
  try {
    dataprep_fields(C->prep, z, zvtx, zeta, cd, cdvtx, nullptr, nullptr, nullptr, nullptr);
  } catch (const std::bad_alloc&) {
    return set_err(C, STORM_E_NOMEM, "out of host memory");
  }
  return STORM_OK;
}

int storm_fvt_get_output(storm_fvt_t* C, double* cdvtx, double* e_est) {
  if (!C) return STORM_E_ARG;
  if (C->vsrc != storm_fvt::V_FINAL)
    return set_err(C, STORM_E_STATE, "the post-loop fields are formed by storm_fvt_finalize");
  try {
    HIPCHK(hipStreamSynchronize(C->stream));
    if (cdvtx) copy_vertices(C, C->d_cdvtx, cdvtx);
    if (e_est) copy_cells(C, C->d_eest, e_est);
  } catch (const StormErr& e) {
    return set_err(C, e.code, e.msg);
  }
  return STORM_OK;
}

int64_t storm_fvt_get_c2v_weights(storm_fvt_t* C, int32_t* ptr, int32_t* cells, double* w,
                                  double* wid) {
  if (!C) return STORM_E_ARG;
  const Prep& P = C->prep;
  const size_t PP = (size_t)P.P;
  const int64_t total = P.vptr[P.P];
  if (!ptr && !cells && !w && !wid) return total;
  // input vertex p's row starts where the rows of vertices 0..p-1 end
  std::vector<int64_t> start(P.P + 1, 0);
  for (int p = 0; p < P.P; ++p) {
    const int pp = P.viperm[p];
    start[p + 1] = start[p] + (P.vptr[pp + 1] - P.vptr[pp]);
  }
  for (int p = 0; p < P.P; ++p) {
    const int pp = P.viperm[p];
    const int deg = P.vptr[pp + 1] - P.vptr[pp];
    if (ptr) ptr[p] = (int32_t)start[p];
    double tdist = 0.0;
    for (int t = 0; t < deg; ++t) tdist = tdist + P.vwt[(size_t)t * PP + pp];
    for (int t = 0; t < deg; ++t) {
      const int64_t at = start[p] + t;
      const double wi = P.vwt[(size_t)t * PP + pp];
      if (cells) cells[at] = P.perm[P.vell[(size_t)t * PP + pp]];
      if (w) w[at] = wi / tdist;
      if (wid) wid[at] = wi;
    }
  }
  if (ptr) ptr[P.P] = (int32_t)start[P.P];
  return total;
}

int storm_fvt_get_gradients(storm_fvt_t* C, double* dhx, double* dhy) {
  if (!C) return STORM_E_ARG;
  try {
    if (!C->started) start_run(C);
    Prep& P = C->prep;
    const size_t N = P.N;
    if (C->iter_host == 0) {  // no stage has run: cell_grads has not either
      for (size_t i = 0; i < N; ++i) {
        if (dhx) dhx[i] = 0.0;
        if (dhy) dhy[i] = 0.0;
      }
      return STORM_OK;
    }
    double* d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(double) * 2 * N));
    const int bs = 256;
    // the first Green-Gauss pass of that stage, again: steps launched after a
    // steady stop run k_gauss1 on inputs that no stage used
    if (C->o.gradient == STORM_GRAD_GAUSS && C->o.limiter != 0)
      hipLaunchKernelGGL(k_gauss1, dim3((P.N + bs - 1) / bs), dim3(bs), 0, C->stream, C->g,
                         C->last_din, C->S.gwork);
    hipLaunchKernelGGL(k_delh, dim3((P.N + bs - 1) / bs), dim3(bs), 0, C->stream, C->g, C->o,
                       C->last_din, C->S.gwork, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(C->stream);
    if (e == hipSuccess) {
      if (dhx) e = hipMemcpy(dhx, d, sizeof(double) * N, hipMemcpyDeviceToHost);
      if (e == hipSuccess && dhy)
        e = hipMemcpy(dhy, d + N, sizeof(double) * N, hipMemcpyDeviceToHost);
    }
    (void)hipFree(d);
    HIPCHK(e);
    std::vector<double> tmp(N);
    for (double* a : {dhx, dhy}) {
      if (!a) continue;
      std::memcpy(tmp.data(), a, sizeof(double) * N);
      for (size_t ii = 0; ii < N; ++ii) a[P.perm[ii]] = tmp[ii];
    }
  } catch (const StormErr& e) {
    return set_err(C, e.code, e.msg);
  }
  return STORM_OK;
}

int storm_fvt_get_gradients_noalloc(storm_fvt_t* C, double* dhx, double* dhy) {
  if (!C) return STORM_E_ARG;
  const Prep& P = C->prep;
  const size_t N = P.N;
  if (P.N > NOALLOC_MAX_CELLS)
    return set_err(C, STORM_E_ARG, "get_gradients_noalloc: more cells than its scratch holds");
  if (!C->started) return set_err(C, STORM_E_STATE, "get_gradients_noalloc: no run started");
  if (C->iter_host == 0) {  // no stage has run: cell_grads has not either
    if (dhx) std::memset(dhx, 0, sizeof(double) * N);
    if (dhy) std::memset(dhy, 0, sizeof(double) * N);
    return STORM_OK;
  }
  const int bs = 256;
  const dim3 gc((P.N + bs - 1) / bs);
  if (C->o.gradient == STORM_GRAD_GAUSS && C->o.limiter != 0)
    hipLaunchKernelGGL(k_gauss1, gc, dim3(bs), 0, C->stream, C->g, C->last_din, C->S.gwork);
  hipLaunchKernelGGL(k_delh_input_order, gc, dim3(bs), 0, C->stream, C->g, C->o, C->last_din,
                     C->S.gwork, (const int*)C->d_perm, C->scratch_ptr);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && dhx)
    e = hipMemcpyAsync(dhx, C->scratch_ptr, sizeof(double) * N, hipMemcpyDeviceToHost, C->stream);
  if (e == hipSuccess && dhy)
    e = hipMemcpyAsync(dhy, C->scratch_ptr + N, sizeof(double) * N, hipMemcpyDeviceToHost,
                       C->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(C->stream);
  if (e != hipSuccess) return set_err(C, STORM_E_DEVICE, "get_gradients_noalloc: HIP error");
  return STORM_OK;
}

int storm_fvt_drynwet(storm_fvt_t* C, const double* z, const double* h, double* zout,
                      int32_t* pd) {
  if (!C || !z || !h || !zout) return set_err(C, STORM_E_ARG, "null argument");
  try {
    const Prep& P = C->prep;
    CallBufs B;
    const double* dz = vertex_upload(C, B, z);
    const double* dh = vertex_upload(C, B, h);
    double* dzo = B.get<double>(3 * (size_t)P.N);
    int* dpd = pd ? B.get<int>(4 * (size_t)P.N) : nullptr;
    const int bs = 256;
    hipLaunchKernelGGL(k_drynwet, dim3((P.N + bs - 1) / bs), dim3(bs), 0, C->stream, C->g,
                       C->o.h_dry, dz, dh, (const int*)C->d_perm, dzo, dpd);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(C->stream));
    HIPCHK(hipMemcpy(zout, dzo, sizeof(double) * 3 * (size_t)P.N, hipMemcpyDeviceToHost));
    if (pd) HIPCHK(hipMemcpy(pd, dpd, sizeof(int) * 4 * (size_t)P.N, hipMemcpyDeviceToHost));
  } catch (const StormErr& e) {
    return set_err(C, e.code, e.msg);
  } catch (const std::bad_alloc&) {
    return set_err(C, STORM_E_NOMEM, "out of host memory");
  }
  return STORM_OK;
}

int storm_fvt_fvsources(storm_fvt_t* C, const double* z, const double* h, double* source) {
  if (!C || !z || !h || !source) return set_err(C, STORM_E_ARG, "null argument");
  try {
    const Prep& P = C->prep;
    if (!C->d_vperm) {
      C->d_vperm = dupload(C->owned, P.vperm);
      C->d_nn_ptr = dupload(C->owned, P.nn_ptr);
      C->d_nn_idx = dupload(C->owned, P.nn_idx);
    }
    CallBufs B;
    const double* dz = vertex_upload(C, B, z);
    const double* dh = vertex_upload(C, B, h);
    double* zb = B.get<double>(2 * (size_t)P.P);
    double* src = B.get<double>(3 * (size_t)P.P);
    const int bs = 256;
    const dim3 gv((P.P + bs - 1) / bs);
    hipLaunchKernelGGL(k_node_grad, gv, dim3(bs), 0, C->stream, C->g, (const int*)C->d_nn_ptr,
                       (const int*)C->d_nn_idx, dz, dh, zb);
    hipLaunchKernelGGL(k_node_integral, gv, dim3(bs), 0, C->stream, C->g, (const double*)zb,
                       (const int*)C->d_vperm, src);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(C->stream));
    HIPCHK(hipMemcpy(source, src, sizeof(double) * 3 * (size_t)P.P, hipMemcpyDeviceToHost));
  } catch (const StormErr& e) {
    return set_err(C, e.code, e.msg);
  } catch (const std::bad_alloc&) {
    return set_err(C, STORM_E_NOMEM, "out of host memory");
  }
  return STORM_OK;
}

// A captured step graph holds the kernel arguments of its capture: the
// source-term tables set after it need a new capture.
static void drop_graph(storm_fvt* C) {
  if (C->gexec) {
    (void)hipGraphExecDestroy(C->gexec);
    C->gexec = nullptr;
  }
}

int storm_fvt_set_wind(storm_fvt_t* C, int64_t n, const double* c, const double* wx,
                       const double* wy) {
  if (!C) return STORM_E_ARG;
  if (n < 0) return set_err(C, STORM_E_ARG, "negative wind data count");
  try {
    const Prep& P = C->prep;
    HIPCHK(hipStreamSynchronize(C->stream));  // a step may still read the old values
    drop_graph(C);
    if (n == 0) {
      C->o.wind = 0;
      return STORM_OK;
    }
    std::vector<double> w;
    std::string err;
    const int rc = wind_coefs(P, n, c, wx, wy, w, err);
    if (rc != STORM_OK) return set_err(C, rc, err);
    if (!C->d_wind) C->d_wind = reinterpret_cast<double2*>(dalloc<double>(C->owned, 2 * (size_t)P.N));
    HIPCHK(hipMemcpy(C->d_wind, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
    C->g.wind = C->d_wind;
    C->o.wind = 1;
  } catch (const StormErr& e) {
    return set_err(C, e.code, e.msg);
  } catch (const std::bad_alloc&) {
    return set_err(C, STORM_E_NOMEM, "out of host memory");
  }
  return STORM_OK;
}

int storm_fvt_set_culverts(storm_fvt_t* C, int32_t n, const double* pin, const double* pout,
                           const int32_t* ntab, const double* head, const double* q) {
  if (!C) return STORM_E_ARG;
  if (n < 0) return set_err(C, STORM_E_ARG, "negative culvert count");
  if (n > 0 && (!pin || !pout || !ntab || !head || !q))
    return set_err(C, STORM_E_ARG, "null culvert array");
  try {
    Prep& P = C->prep;
    if (n > 0 && P.nr > 1)
      return set_err(C, STORM_E_ARG, "culverts need a context of one reach");
    std::vector<int> tptr(n + 1, 0);
    for (int m = 0; m < n; ++m) {
      if (ntab[m] < 1) return set_err(C, STORM_E_ARG, "a culvert curve needs at least one point");
      tptr[m + 1] = tptr[m] + ntab[m];
    }
    std::vector<int> cin, cout;
    std::vector<double> rin;
    std::string err;
    if (n > 0) {
      const int rc = locate_culverts(P, n, pin, pout, cin, cout, rin, err);
      if (rc != STORM_OK) return set_err(C, rc, err);
    }
    HIPCHK(hipStreamSynchronize(C->stream));  // a step may still read the old tables
    drop_graph(C);
    for (void* b : C->cv_owned) (void)hipFree(b);
    C->cv_owned.clear();
    C->cv = CvTab{};
    C->d_cvsrc = nullptr;
    C->g.cv_io = nullptr;
    C->g.cvsrc = nullptr;
    C->o.ncv = 0;
    const uint64_t flag = 1ull << CM_CULVERT;
    for (int i = 0; i < P.N; ++i) P.cmeta[i] &= ~flag;
    std::vector<int> io(2 * (size_t)n);
    std::vector<double> zinv(n);
    for (int m = 0; m < n; ++m) {
      io[2 * m] = P.iperm[cin[m]];
      io[2 * m + 1] = P.iperm[cout[m]];
      P.cmeta[io[2 * m]] |= flag;
      P.cmeta[io[2 * m + 1]] |= flag;
      zinv[m] = pin[3 * m + 2];
    }
    HIPCHK(hipMemcpy(C->d_cmeta, P.cmeta.data(), sizeof(uint64_t) * P.N, hipMemcpyHostToDevice));
    C->cv_in = cin;
    C->cv_out = cout;
    if (n == 0) return STORM_OK;
    auto& own = C->cv_owned;
    C->g.cv_io = reinterpret_cast<const int2*>(dupload(own, io));
    C->cv.rin = reinterpret_cast<const double2*>(dupload(own, rin));
    C->cv.zinv = dupload(own, zinv);
    C->cv.tptr = dupload(own, tptr);
    C->cv.head = dupload(own, std::vector<double>(head, head + tptr[n]));
    C->cv.q = dupload(own, std::vector<double>(q, q + tptr[n]));
    C->d_cvsrc = dalloc<double>(own, n);
    C->g.cvsrc = C->d_cvsrc;
    C->o.ncv = n;
  } catch (const StormErr& e) {
    return set_err(C, e.code, e.msg);
  } catch (const std::bad_alloc&) {
    return set_err(C, STORM_E_NOMEM, "out of host memory");
  }
  return STORM_OK;
}

int storm_fvt_get_culverts(storm_fvt_t* C, int32_t* cell_in, int32_t* cell_out, double* src) {
  if (!C) return STORM_E_ARG;
  try {
    const int n = C->o.ncv;
    for (int m = 0; m < n; ++m) {
      if (cell_in) cell_in[m] = C->cv_in[m];
      if (cell_out) cell_out[m] = C->cv_out[m];
    }
    if (src && n > 0) {
      HIPCHK(hipStreamSynchronize(C->stream));
      HIPCHK(hipMemcpy(src, C->d_cvsrc, sizeof(double) * n, hipMemcpyDeviceToHost));
    }
  } catch (const StormErr& e) {
    return set_err(C, e.code, e.msg);
  }
  return STORM_OK;
}

int storm_fvt_get_bc_values(storm_fvt_t* C, double* qin, double* hbc) {
  if (!C) return STORM_E_ARG;
  try {
    if (!C->started) start_run(C);
    HIPCHK(hipStreamSynchronize(C->stream));
    Prep& P = C->prep;
    if (qin && P.n_in)
      HIPCHK(hipMemcpy(qin, C->S.sq, sizeof(double) * P.n_in, hipMemcpyDeviceToHost));
    if (hbc && P.n_out)
      HIPCHK(hipMemcpy(hbc, C->S.sh, sizeof(double) * P.n_out, hipMemcpyDeviceToHost));
  } catch (const StormErr& e) {
    return set_err(C, e.code, e.msg);
  }
  return STORM_OK;
}

int storm_fvt_device_info(storm_fvt_t* C, int32_t* n_elems, int32_t* perm, void** stream) {
  if (!C) return STORM_E_ARG;
  if (n_elems) *n_elems = C->prep.N;
  if (perm) std::memcpy(perm, C->prep.perm.data(), sizeof(int32_t) * C->prep.N);
  if (stream) *stream = (void*)C->stream;
  return STORM_OK;
}

double storm_fvt_last_run_seconds(storm_fvt_t* C) { return C ? C->last_run_s : -1.0; }

const char* storm_fvt_last_error(storm_fvt_t* C) { return C ? C->err.c_str() : "null context"; }

void storm_fvt_destroy(storm_fvt_t* C) { delete C; }

int64_t storm_fvt_inspect_molecule(const storm_mesh_t* mesh, int32_t* ptr, int32_t* cells) {
  if (!mesh) return STORM_E_ARG;
  std::vector<int> p, c;
  std::string err;
  int rc = build_molecule(mesh, p, c, err);
  if (rc != STORM_OK) return rc;
  if (ptr) std::memcpy(ptr, p.data(), sizeof(int32_t) * p.size());
  if (cells && !c.empty()) std::memcpy(cells, c.data(), sizeof(int32_t) * c.size());
  return (int64_t)c.size();
}

}  // extern "C"
