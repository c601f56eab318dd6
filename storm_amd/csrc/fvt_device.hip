// MI355X (gfx950) FP64 kernels of the SToRM FVT time step and the C ABI
// declared in include/storm_fvt.h.
//
// Per RK stage (solverFVT.f90:158-331) four launches:
//   k_vertex  to_vertices' vertex pass (to_vertices.f90:26-41, enforce_bcs
//             stage nodes): IDW water surface over wet cells -> hvtx + wet flag
//   k_grad    find_dry_cells' results + cell_grads (LSQ/Gauss, Barth-Jespersen)
//             + hgradfix2 + the cell pass of to_vertices (:44-55)
//   k_bcq     inflowQbyH's fixed point, one block per discharge string
//   k_update  every flux of the cell's three edges, gathered by the cell in
//             the reference's accumulation order (no atomics), bed slope,
//             explicit friction if chosen, SSP-RK update, clip and wet/dry
//             recovery; in the last stage also implicit friction, hvalve,
//             the residual partial sums and the next step's CFL candidate
// and per step one single-block k_control (dt, time, BC tables, stopping).
//
// Cells are in Reverse Cuthill-McKee order; all arrays are structure-of-
// arrays FP64/int32; per-cell edge data are stored per (slot, cell).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/storm_fvt.h"
#include "fvt_math.h"
#include "fvt_prep.h"

using namespace storm;

namespace {

constexpr int CTRL_THREADS = 256;

struct Ctrl {
  double dt;          // dt of the current step
  double delt_last;   // steady: local dt of the reference's last cell
  double ttime, stime, rmax, sumtop, sumbot;
  int iter, done, pending, pad;
  double qin[MAX_STRINGS], hbc[MAX_STRINGS];
  double atotal[MAX_STRINGS];
  int null_area[MAX_STRINGS], qskip[MAX_STRINGS];
};

struct Geo {
  int N, P, R;
  const int *cvtx, *cnbr, *ep1, *ep2;
  const uint32_t* cinfo;
  const uint8_t* corder;
  const double *rc, *enx, *eny, *etx, *ety, *elen, *ezmid;
  const double *area, *perim, *z, *cd, *delzx, *delzy, *zhigh, *vdx, *vdy, *zv3;
  const double *xv, *yv;
  const int *mptr, *mcell;
  const double *mdx, *mdy, *mw, *ai, *bi, *ci, *det;
  const int* gconn;
  const double *gcx, *gcy, *gwt;
  const int *vptr, *vcell, *vbc;
  const double *vw, *zvtx;
  const int *qs_ptr, *qs_cell, *qs_slot;
  const int* in_type;
  int n_in, n_out, last_cell;
  double bch_max;
};

struct Opts {
  int rkorder, dt_op, steady, stop_cr, btype, hvalve, limiter, flux, gradient, friction, vclip;
  double cfl, rstop, h_dry, h_wet, vceiling;
};

struct Stage {
  const double* slot[4];  // rk slots 0..s-1 (h at +0, hu at +N, hv at +2N)
  double alpha[4];
  double* out;            // slot s
  const double* hin;      // h of the stage input (slot s-1)
  double beta;
  int s, last;
};

struct State {
  double *u, *v, *up, *vp, *grad, *gwork, *hvtx, *zetavtx, *phi, *part;
  uint8_t *dry_in, *dry_out, *dry_post, *vwet;
  Ctrl* ctrl;
};

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
__global__ void __launch_bounds__(256)
k_init(Geo g, Opts o, double* slot0, const double* h0, const double* u0, const double* v0,
       double* u, double* v, uint8_t* dry, double* part) {
  // solverFVT.f90:123-127 and the first find_dry_cells; the CFL candidates of
  // step 1 go to the per-block partials (slot 3), reduced by k_control
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double cand = DBL_MAX;
  if (i < g.N) {
    const double h = h0[i], uu = u0[i], vv = v0[i];
    slot0[g.N + i] = h * uu;
    slot0[2 * g.N + i] = h * vv;
    slot0[i] = h;
    u[i] = uu;
    v[i] = vv;
    double um;
    bool d = find_dry(h, uu, vv, o.h_dry, o.h_wet, &um);
    dry[i] = d;
    const double wc = d ? ZERO : sqrt(G * h);
    double dd = um + wc + SMALL;
    cand = o.cfl * g.area[i] / (dd * g.perim[i]);
  }
  __shared__ double s_c[256];
  const int t = threadIdx.x;
  s_c[t] = cand;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) s_c[t] = dmin(s_c[t], s_c[t + w]);
    __syncthreads();
  }
  if (t == 0) {
    part[4 * blockIdx.x] = ZERO;
    part[4 * blockIdx.x + 1] = ZERO;
    part[4 * blockIdx.x + 2] = ZERO;
    part[4 * blockIdx.x + 3] = s_c[0];
  }
}

// ---------------------------------------------------------------- control
__device__ void get_time_bcs(const Geo& g, const int* out_type, const int* qts_ptr,
                             const double* qts_t, const double* qts_v, const int* hts_ptr,
                             const double* hts_t, const double* hts_v, double time, Ctrl* c) {
  // get_time_bcs.f90:31-77
  for (int k = 0; k < g.n_out; ++k) {
    if (out_type[k] != 1) continue;
    int a = hts_ptr[k], n = hts_ptr[k + 1] - a;
    if (n <= 0) continue;
    const double* tl = hts_t + a;
    const double* tv = hts_v + a;
    if (time < tl[0]) c->hbc[k] = tv[0];
    else if (time >= tl[n - 1]) c->hbc[k] = tv[n - 1];
    else
      for (int i = 0; i < n - 1; ++i)
        if (time >= tl[i] && time < tl[i + 1]) {
          double d = time - tl[i], d2 = tl[i + 1] - tl[i];
          c->hbc[k] = tv[i] + d * (tv[i + 1] - tv[i]) / d2;
          break;
        }
  }
  for (int k = 0; k < g.n_in; ++k) {
    int a = qts_ptr[k], n = qts_ptr[k + 1] - a;
    if (n <= 0) continue;
    const double* tl = qts_t + a;
    const double* tv = qts_v + a;
    if (time < tl[0]) c->qin[k] = tv[0];
    else if (time >= tl[n - 1]) c->qin[k] = tv[n - 1];
    else
      for (int i = 0; i < n - 1; ++i)
        if (time >= tl[i] && time < tl[i + 1]) {
          double d = time - tl[i], d2 = tl[i + 1] - tl[i];
          c->qin[k] = tv[i] + d * (tv[i + 1] - tv[i]) / d2;
          break;
        }
  }
}

// Step bookkeeping of solverFVT.f90:132-155 (begin) and :338-472 (end of the
// previous step: residuals, stopping test, time, boundary tables).
__global__ void __launch_bounds__(CTRL_THREADS)
k_control(Geo g, Opts o, Ctrl* c, const double* part, int nparts, int begin, const int* out_type,
          const int* qts_ptr, const double* qts_t, const double* qts_v, const int* hts_ptr,
          const double* hts_t, const double* hts_v) {
  __shared__ double s_top[CTRL_THREADS], s_bot[CTRL_THREADS], s_max[CTRL_THREADS],
      s_dt[CTRL_THREADS];
  const int t = threadIdx.x;
  if (c->done) return;
  // fixed-order reduction of the per-block partials of the last stage
  // (residual sums, max|phi|) and of the next step's CFL candidates
  double top = ZERO, bot = ZERO, mx = ZERO, dtm = DBL_MAX;
  for (int b = t; b < nparts; b += CTRL_THREADS) {
    top = top + part[4 * b];
    bot = bot + part[4 * b + 1];
    mx = dmax(mx, part[4 * b + 2]);
    dtm = dmin(dtm, part[4 * b + 3]);
  }
  s_top[t] = top; s_bot[t] = bot; s_max[t] = mx; s_dt[t] = dtm;
  __syncthreads();
  for (int w = CTRL_THREADS / 2; w > 0; w >>= 1) {
    if (t < w) {
      s_top[t] = s_top[t] + s_top[t + w];
      s_bot[t] = s_bot[t] + s_bot[t + w];
      s_max[t] = dmax(s_max[t], s_max[t + w]);
      s_dt[t] = dmin(s_dt[t], s_dt[t + w]);
    }
    __syncthreads();
  }
  if (t != 0) return;
  if (c->pending) {
    double sumtop = s_top[0], sumbot = s_bot[0];
    c->sumtop = sumtop;
    c->sumbot = sumbot;
    if (o.steady) {
      if (o.stop_cr == 0) c->rmax = s_max[0];
      else {
        if (equals(sumbot, ZERO)) sumbot = ONE;
        c->rmax = sqrt(sumtop / sumbot);
      }
      c->stime = (double)c->iter;
      c->pending = 0;
      if (c->rmax < o.rstop) { c->done = 1; return; }
    } else {
      c->stime = c->stime + c->dt;
      c->rmax = s_max[0];
      get_time_bcs(g, out_type, qts_ptr, qts_t, qts_v, hts_ptr, hts_t, hts_v, c->stime, c);
      c->pending = 0;
    }
  }
  if (begin) {
    c->iter = c->iter + 1;
    if (!o.steady) {
      c->dt = (o.dt_op == 0) ? o.cfl : s_dt[0];
      c->ttime = c->ttime + c->dt;
    } else {
      c->ttime = c->ttime + c->delt_last;
    }
    c->pending = 1;
  }
}

// ---------------------------------------------------------------- vertex pass
// to_vertices.f90:26-41 with ctr2vtxWSE.f90:89-105 and positivity_vtx.f90,
// then the stage strings' nodes of enforce_bcs.f90:40-56.
__global__ void k_vertex(Geo g, Opts o, const double* hin, const uint8_t* dry, double* hvtx,
                         uint8_t* vwet, double* zetavtx, const Ctrl* c) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.P || c->done) return;
  const double zv = g.zvtx[p];
  double id = ZERO, a = ZERO, zr = zv;
  int counter = 0;
  for (int t = g.vptr[p]; t < g.vptr[p + 1]; ++t) {
    const int k = g.vcell[t];
    if (dry[k]) continue;
    counter++;
    const double w = g.vw[t];
    const double zk = hin[k] + g.z[k];
    id = id + w;
    a = a + w * zk;
    zr = dmax(zk, zr);
  }
  double zeta = (counter > 0) ? a / id : zv;
  zeta = dmin(zeta, zr);
  double hv = zeta - zv;
  if (hv < ZERO) { hv = ZERO; zeta = zv; }
  vwet[p] = hv > o.h_dry;
  const int b = g.vbc[p];
  if (b >= 0) hv = dmax(ZERO, c->hbc[b] - zv);
  hvtx[p] = hv;
  if (zetavtx) zetavtx[p] = zeta;
}

// ---------------------------------------------------------------- Green-Gauss pass 1
__global__ void k_gauss1(Geo g, const double* hin, const double* u, const double* v,
                         double* gwork, const Ctrl* c) {
  // gaussgrad.f90:37-42 for zeta, u, v
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.N || c->done) return;
  const int N = g.N;
  const int n0 = g.gconn[i], n1 = g.gconn[N + i], n2 = g.gconn[2 * N + i];
  const double cx0 = g.gcx[i], cx1 = g.gcx[N + i], cx2 = g.gcx[2 * N + i];
  const double cy0 = g.gcy[i], cy1 = g.gcy[N + i], cy2 = g.gcy[2 * N + i];
  const double z0 = hin[n0] + g.z[n0], z1 = hin[n1] + g.z[n1], z2 = hin[n2] + g.z[n2];
  gwork[i] = cx0 * z0 + cx1 * z1 + cx2 * z2;
  gwork[N + i] = cy0 * z0 + cy1 * z1 + cy2 * z2;
  gwork[2 * N + i] = cx0 * u[n0] + cx1 * u[n1] + cx2 * u[n2];
  gwork[3 * N + i] = cy0 * u[n0] + cy1 * u[n1] + cy2 * u[n2];
  gwork[4 * N + i] = cx0 * v[n0] + cx1 * v[n1] + cx2 * v[n2];
  gwork[5 * N + i] = cy0 * v[n0] + cy1 * v[n1] + cy2 * v[n2];
}

// limiterBJ.f90:42-81, beta = 1 (cell_grads.f90:59)
__device__ __forceinline__ void bj_limit(double q0, const double qn[3], const double rcx[3],
                                         const double rcy[3], double& gx, double& gy) {
  double r[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double qmin = dmin(q0, qn[j]), qmax = dmax(q0, qn[j]);
    const double qj = q0 + gx * rcx[j] + gy * rcy[j];
    const double dq = qj - q0;
    double rr;
    if (equals(dq, ZERO)) rr = ONE;
    else if (dq < 0) rr = (qmin - q0) / dq;
    else rr = (qmax - q0) / dq;
    const double a = dmin(ONE * rr, ONE), b = dmin(rr, ONE);
    r[j] = dmax(a, b);
  }
  const double ph = dmin(dmin(r[0], r[1]), r[2]);
  gx = ph * gx;
  gy = ph * gy;
}

// ---------------------------------------------------------------- gradients
// cell_grads.f90:24-75, hgradfix2.f90:46-98 and to_vertices.f90:44-55.
__global__ void k_grad(Geo g, Opts o, Stage st, State S) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.N || S.ctrl->done) return;
  const int N = g.N;
  const double* h = st.hin;
  const double hi = h[i], zi = g.z[i];
  const double zeta_i = hi + zi;
  const double ui = S.u[i], vi = S.v[i];
  const bool dry_i = S.dry_in[i];
  double gzx = ZERO, gzy = ZERO, gux = ZERO, guy = ZERO, gvx = ZERO, gvy = ZERO;
  if (o.limiter != 0) {
    if (o.gradient == STORM_GRAD_GAUSS) {
      // gaussgrad.f90:45-52
      const double* w = S.gwork;
      const int n0 = g.gconn[i], n1 = g.gconn[N + i], n2 = g.gconn[2 * N + i];
      const double w0 = g.gwt[i], w1 = g.gwt[N + i], w2 = g.gwt[2 * N + i];
      gzx = HALF * (w[i] + w0 * w[n0] + w1 * w[n1] + w2 * w[n2]);
      gzy = HALF * (w[N + i] + w0 * w[N + n0] + w1 * w[N + n1] + w2 * w[N + n2]);
      gux = HALF * (w[2 * N + i] + w0 * w[2 * N + n0] + w1 * w[2 * N + n1] + w2 * w[2 * N + n2]);
      guy = HALF * (w[3 * N + i] + w0 * w[3 * N + n0] + w1 * w[3 * N + n1] + w2 * w[3 * N + n2]);
      gvx = HALF * (w[4 * N + i] + w0 * w[4 * N + n0] + w1 * w[4 * N + n1] + w2 * w[4 * N + n2]);
      gvy = HALF * (w[5 * N + i] + w0 * w[5 * N + n0] + w1 * w[5 * N + n1] + w2 * w[5 * N + n2]);
    } else {
      // lsq_grad_array.f90:60-73 for zeta, u, v
      double dz = ZERO, ez = ZERO, du = ZERO, eu = ZERO, dv = ZERO, ev = ZERO;
      for (int t = g.mptr[i]; t < g.mptr[i + 1]; ++t) {
        const int k = g.mcell[t];
        const double w = g.mw[t], dx = g.mdx[t], dy = g.mdy[t];
        const double fz = (h[k] + g.z[k]) - zeta_i;
        const double fu = S.u[k] - ui;
        const double fv = S.v[k] - vi;
        dz = dz + w * fz * dx; ez = ez + w * fz * dy;
        du = du + w * fu * dx; eu = eu + w * fu * dy;
        dv = dv + w * fv * dx; ev = ev + w * fv * dy;
      }
      const double a = g.ai[i], b = g.bi[i], cc = g.ci[i], d = g.det[i];
      gzx = (cc * dz - b * ez) / d; gzy = (a * ez - dz * b) / d;
      gux = (cc * du - b * eu) / d; guy = (a * eu - du * b) / d;
      gvx = (cc * dv - b * ev) / d; gvy = (a * ev - dv * b) / d;
    }
    if (o.limiter == 2) {
      double qz[3], qu[3], qv[3], rcx[3], rcy[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int k = g.cnbr[j * N + i];
        qz[j] = (k < 0) ? zeta_i : h[k] + g.z[k];
        qu[j] = (k < 0) ? ui : S.u[k];
        qv[j] = (k < 0) ? vi : S.v[k];
        rcx[j] = g.rc[(2 * j) * N + i];
        rcy[j] = g.rc[(2 * j + 1) * N + i];
      }
      bj_limit(zeta_i, qz, rcx, rcy, gzx, gzy);
      bj_limit(ui, qu, rcx, rcy, gux, guy);
      bj_limit(vi, qv, rcx, rcy, gvx, gvy);
    }
  }
  const double dzx = g.delzx[i], dzy = g.delzy[i];
  double dhx = gzx - dzx, dhy = gzy - dzy;
  double up = ui, vp = vi;
  if (!dry_i) {
    // hgradfix2.f90:49-96
    bool setgrad = zeta_i < g.zhigh[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int l = g.cnbr[k * N + i];
      if (l < 0) continue;
      if (S.dry_in[l] && g.z[l] < zeta_i) setgrad = false;
    }
    if (setgrad) {
      dhx = -dzx; dhy = -dzy;
      up = ZERO; vp = ZERO;
      double* sl = const_cast<double*>(st.hin);  // slot s-1: zero its hu, hv
      sl[N + i] = ZERO;
      sl[2 * N + i] = ZERO;
      gux = ZERO; guy = ZERO; gvx = ZERO; gvy = ZERO;
    }
    // to_vertices.f90:44-55
    const bool flag = S.vwet[g.cvtx[i]] && S.vwet[g.cvtx[N + i]] && S.vwet[g.cvtx[2 * N + i]];
    if (!flag) { up = ZERO; vp = ZERO; }
  }
  S.grad[i] = dhx; S.grad[N + i] = dhy;
  S.grad[2 * N + i] = gux; S.grad[3 * N + i] = guy;
  S.grad[4 * N + i] = gvx; S.grad[5 * N + i] = gvy;
  S.up[i] = up;
  S.vp[i] = vp;
  if (o.steady && st.last && i == g.last_cell) {
    // the local dt of the reference's last cell, which fric_terms_impl uses
    // (solverFVT.f90:260-265,336)
    double umag = ZERO;
    bool d = find_dry(hi, ui, vi, o.h_dry, o.h_wet, &umag);
    const double wc = d ? ZERO : sqrt(G * hi);
    double dd = umag + wc;
    if (dd < SMALL) dd = 1.044;
    S.ctrl->delt_last = o.cfl * g.area[i] / (dd * g.perim[i]);
  }
}

// face value of cell c at its local edge m: q + grad(q).rc(c,m)
struct Face { double h, u, v; };
__device__ __forceinline__ Face face(const Geo& g, const State& S, const double* h, int c, int m) {
  const int N = g.N;
  const double rx = g.rc[(2 * m) * N + c], ry = g.rc[(2 * m + 1) * N + c];
  Face f;
  f.h = h[c] + S.grad[c] * rx + S.grad[N + c] * ry;
  f.u = S.up[c] + S.grad[2 * N + c] * rx + S.grad[3 * N + c] * ry;
  f.v = S.vp[c] + S.grad[4 * N + c] * rx + S.grad[5 * N + c] * ry;
  return f;
}

// ---------------------------------------------------------------- inflowQbyH
// inflowQbyH.f90:31-126 (vtype 1) / :221-316 (vtype 5): the fixed point on
// the string's wetted length, one block per string.  Sums run in string
// order (per-edge terms formed in parallel, accumulated by one lane).
constexpr int BCQ_THREADS = 256;
__global__ void __launch_bounds__(BCQ_THREADS)
k_bcq(Geo g, Opts o, Stage st, State S) {
  const int ch = blockIdx.x;
  Ctrl* c = S.ctrl;
  if (c->done) return;
  const int vt = g.in_type[ch];
  if (vt != 1 && vt != 5) return;
  const int t = threadIdx.x;
  const int e0 = g.qs_ptr[ch], e1 = g.qs_ptr[ch + 1];
  const double qin = c->qin[ch];
  if (qin < DBL_MIN) { if (t == 0) c->qskip[ch] = 1; return; }
  __shared__ double s_a[BCQ_THREADS], s_q[BCQ_THREADS];
  __shared__ char s_ok[BCQ_THREADS];
  __shared__ double s_atotal, s_new, s_qt;
  __shared__ int s_null, s_exit;
  const double h_dry = o.h_dry;
  const double* h = st.hin;
  // initial wetted length
  if (t == 0) s_atotal = ZERO;
  __syncthreads();
  for (int base = e0; base < e1; base += BCQ_THREADS) {
    const int l = base + t;
    s_ok[t] = 0;
    if (l < e1) {
      const int i = g.qs_cell[l], m = g.qs_slot[l];
      if (!(vt == 5 && S.dry_in[i])) {
        double hl;
        if (S.dry_in[i]) hl = ZERO;
        else {
          hl = h[i] + S.grad[i] * g.rc[(2 * m) * g.N + i] + S.grad[g.N + i] * g.rc[(2 * m + 1) * g.N + i];
          hl = dmax(hl, ZERO);
        }
        s_a[t] = hl * g.elen[m * g.N + i];
        s_ok[t] = 1;
      }
    }
    __syncthreads();
    if (t == 0) {
      double a = s_atotal;
      for (int k = 0; k < BCQ_THREADS && base + k < e1; ++k)
        if (s_ok[k]) a = a + s_a[k];
      s_atotal = a;
    }
    __syncthreads();
  }
  if (t == 0) { s_null = equals(s_atotal, ZERO); s_exit = 0; }
  __syncthreads();
  int counter = 0;
  for (;;) {
    if (t == 0) { s_new = ZERO; s_qt = ZERO; }
    __syncthreads();
    const double qfrac = qin / s_atotal;
    for (int base = e0; base < e1; base += BCQ_THREADS) {
      const int l = base + t;
      s_ok[t] = 0;
      if (l < e1) {
        const int i = g.qs_cell[l], m = g.qs_slot[l];
        if (!(vt == 5 && S.dry_in[i])) {
          Face f = face(g, S, h, i, m);
          double hl = f.h, ue, ve;
          if (s_null) hl = g.bch_max;
          if (hl < h_dry) { hl = dmax(hl, ZERO); ue = ZERO; ve = ZERO; }
          else { ue = f.u; ve = f.v; }
          const double nx = g.enx[m * g.N + i], ny = g.eny[m * g.N + i], L = g.elen[m * g.N + i];
          double ur = -qfrac;
          const double ul = ue * nx + ve * ny;
          const double sq = sqrt(G * hl);
          double hr = 2.0 * sq + ul - ur;
          hr = hr * hr / 4.0 / G;
          ur = -qfrac * nx;
          const double vr = -qfrac * ny;
          if (!(hr < h_dry)) {
            s_a[t] = hr * L;
            s_q[t] = (ur * nx + vr * ny) * hr * L;
            s_ok[t] = 1;
          }
        }
      }
      __syncthreads();
      if (t == 0) {
        double a = s_new, q = s_qt;
        for (int k = 0; k < BCQ_THREADS && base + k < e1; ++k)
          if (s_ok[k]) { a = a + s_a[k]; q = q + s_q[k]; }
        s_new = a; s_qt = q;
      }
      __syncthreads();
    }
    if (t == 0) {
      const double qt = fabs(s_qt);
      if (fabs(qt - qin) / qin < 0.005 || counter > 50) s_exit = 1;
      else s_atotal = s_new;
    }
    __syncthreads();
    if (s_exit) break;
    counter++;
  }
  if (t == 0) {
    c->atotal[ch] = s_new;
    c->null_area[ch] = s_null;
    c->qskip[ch] = 0;
  }
}

// ---------------------------------------------------------------- update
constexpr int UPD_THREADS = 256;

__global__ void __launch_bounds__(UPD_THREADS)
k_update(Geo g, Opts o, Stage st, State S, double* part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const Ctrl* c = S.ctrl;
  const bool done = c->done;
  const int N = g.N;
  double top = ZERO, bot = ZERO, mxphi = ZERO, cand = DBL_MAX;
  const double h_dry = o.h_dry, h_wet = o.h_wet;
  if (i < N && !done) {
    const double* h = st.hin;
    const bool dry_i = S.dry_in[i];
    const double hi = h[i];
    double p0 = ZERO, p1 = ZERO, p2 = ZERO;
    const uint8_t ord = g.corder[i];
#pragma unroll 1
    for (int tt = 0; tt < 3; ++tt) {
      const int j = (ord >> (2 * tt)) & 3;
      const uint32_t info = g.cinfo[j * N + i];
      const uint32_t kind = info & CI_KIND_MASK;
      const double nx = g.enx[j * N + i], ny = g.eny[j * N + i], L = g.elen[j * N + i];
      if (kind == SK_INTERIOR) {
        // invisc_fluxes.f90:39-166
        const int nb = g.cnbr[j * N + i];
        const bool dry_nb = S.dry_in[nb];
        if (dry_i && dry_nb) continue;
        const bool left = info & CI_LEFT;
        const int nl = (info >> CI_NBRLOC_SHIFT) & 3;
        const int cL = left ? i : nb, mL = left ? j : nl;
        const int cR = left ? nb : i, mR = left ? nl : j;
        const bool dL = left ? dry_i : dry_nb, dR = left ? dry_nb : dry_i;
        double hl, ul, vl, hr, ur, vr;
        if (dL) { hl = ZERO; ul = ZERO; vl = ZERO; }
        else {
          Face f = face(g, S, h, cL, mL);
          hl = f.h; ul = f.u; vl = f.v;
          if (hl < ZERO) hl = ZERO;
        }
        if (dR) { hr = ZERO; ur = ZERO; vr = ZERO; }
        else {
          Face f = face(g, S, h, cR, mR);
          hr = f.h; ur = f.u; vr = f.v;
          if (hr < ZERO) hr = ZERO;
        }
        if (hl < h_dry && hr < h_dry) continue;
        bool set_wall = false;
        if (dL || dR) {
          const double h0 = dmax(hr, hl) + g.ezmid[j * N + i];
          if (dL && g.z[cL] + h_wet > h0) { set_wall = true; hl = hr; }
          if (dR && g.z[cR] + h_wet > h0) { set_wall = true; hr = hl; }
        }
        double f[3];
        if (set_wall) {
          f[0] = ZERO;
          hl = dmax(hl, hr);
          f[1] = HALF * G * hl * hl * nx;
          f[2] = HALF * G * hl * hl * ny;
        } else if (o.flux == STORM_FLUX_AC) {
          flux_ac(nx, ny, hl, ul, vl, hr, ur, vr, f);
        } else if (o.flux == STORM_FLUX_RUSANOV) {
          const double wl = dL ? ZERO : sqrt(G * h[cL]), wr = dR ? ZERO : sqrt(G * h[cR]);
          flux_rusanov(nx, ny, hl, ul, vl, wl, hr, ur, vr, wr, f);
        } else if (o.flux == STORM_FLUX_HLL) {
          flux_hll(nx, ny, hl, ul, vl, hr, ur, vr, h_dry, f);
        } else {
          flux_bgnvc(nx, ny, hl, ul, vl, hr, ur, vr, f);
        }
        if (left) { p0 = p0 - f[0] * L; p1 = p1 - f[1] * L; p2 = p2 - f[2] * L; }
        else { p0 = p0 + f[0] * L; p1 = p1 + f[1] * L; p2 = p2 + f[2] * L; }
      } else if (kind == SK_WALL) {
        if (dry_i) continue;
        if (o.btype == 0) {
          // wenslp3_fluxes.f90:40-61
          const double h1 = S.hvtx[g.ep1[j * N + i]], h2 = S.hvtx[g.ep2[j * N + i]];
          double he = sqrt((h1 * h1 + h1 * h2 + h2 * h2) * ONE_THIRD);
          if (he < ZERO) he = ZERO;
          if (he < h_dry) continue;
          p1 = p1 - HALF * G * he * he * nx * L;
          p2 = p2 - HALF * G * he * he * ny * L;
        } else {
          // weslp2_fluxes.f90:40-75
          Face f = face(g, S, h, i, j);
          if (f.h < h_dry) continue;
          const double tx = g.etx[j * N + i], ty = g.ety[j * N + i];
          double urr = f.u * tx + f.v * ty;
          urr = urr + 2.0 * sqrt(G * f.h);
          const double hrr = urr * urr / 4.0 / G;
          const double f1 = HALF * G * hrr * hrr * nx, f2 = HALF * G * hrr * hrr * ny;
          const double dh = ZERO;
          p1 = p1 - (f1 + dh * nx) * L;
          p2 = p2 - (f2 + dh * ny) * L;
        }
      } else if (kind == SK_IN_Q1 || kind == SK_IN_Q5) {
        // inflowQbyH.f90:131-210 (vtype 1) / :321-400 (vtype 5)
        const int ch = (info >> CI_CHAIN_SHIFT) & 0xFF;
        if (c->qskip[ch]) continue;
        if (kind == SK_IN_Q5 && dry_i) continue;
        const double qin = c->qin[ch];
        const double atotal = c->atotal[ch];
        Face fc = face(g, S, h, i, j);
        double hl = fc.h, ue, ve;
        if (c->null_area[ch]) hl = g.bch_max;
        if (hl < h_dry) { hl = dmax(hl, ZERO); ue = ZERO; ve = ZERO; }
        else { ue = fc.u; ve = fc.v; }
        const double qfrac = qin / atotal;
        double urr = -qfrac;
        const double ul = ue * nx + ve * ny;
        const double sq = sqrt(G * hl);
        double hr = 2.0 * sq + ul - urr;
        hr = hr * hr / 4.0 / G;
        urr = -qfrac * nx;
        const double vrr = -qfrac * ny;
        if (hr < h_dry) continue;
        const double f0 = hr * (urr * nx + vrr * ny);
        const double f1 = hr * urr * (urr * nx + vrr * ny) + HALF * G * hr * hr * nx;
        const double f2 = hr * vrr * (urr * nx + vrr * ny) + HALF * G * hr * hr * ny;
        p0 = p0 - f0 * L; p1 = p1 - f1 * L; p2 = p2 - f2 * L;
      } else if (kind == SK_IN_H) {
        // inflowH.f90:31-91
        const int ch = (info >> CI_CHAIN_SHIFT) & 0xFF;
        Face fc = face(g, S, h, i, j);
        double hl = dmax(fc.h, ZERO), ue, ve;
        if (hi < h_dry) { ue = ZERO; ve = ZERO; } else { ue = fc.u; ve = fc.v; }
        const double hr = dmax(ZERO, c->qin[ch] - g.ezmid[j * N + i]);
        const double ul = ue * nx + ve * ny;
        const double sl = sqrt(G * hl), sr = sqrt(G * hr);
        double urr = ul + (sl - sr);
        const double vrr = urr * ny;
        urr = urr * nx;
        if (hl < h_dry && hr < h_dry) continue;
        double f[3];
        flux_ac(nx, ny, hl, ue, ve, hr, urr, vrr, f);
        p0 = p0 - f[0] * L; p1 = p1 - f[1] * L; p2 = p2 - f[2] * L;
      } else if (kind == SK_OUT_H) {
        // outflowbyH.f90:33-135
        const int ch = (info >> CI_CHAIN_SHIFT) & 0xFF;
        Face fc = face(g, S, h, i, j);
        double hl = fc.h, ue, ve;
        if (dry_i || hl < h_dry) { ue = ZERO; ve = ZERO; hl = dmax(hl, ZERO); }
        else { ue = fc.u; ve = fc.v; }
        const double tx = g.etx[j * N + i], ty = g.ety[j * N + i];
        double hr = dmax(ZERO, c->hbc[ch] - g.ezmid[j * N + i]);
        if (hl < h_dry && hr < h_dry) continue;
        double ul, vl, urr, vrr;
        if (hr < h_dry) { urr = ZERO; ul = ue; vrr = ZERO; vl = ve; }
        else {
          ul = ue * nx + ve * ny;
          vl = ue * tx + ve * ty;
          const double sl = sqrt(G * hl), sr = sqrt(G * hr);
          urr = ul + (sl - sr);
          vrr = vl;
          ul = ue;
          vl = ve;
          const double a = urr, b = vrr;
          urr = a * nx + b * tx;
          vrr = a * ny + b * ty;
        }
        bool set_wall = false;
        if (o.hvalve) {
          const double dot = urr * nx + vrr * ny;
          if (dot < ZERO) { set_wall = true; hr = hl; urr = ZERO; vrr = ZERO; }
        }
        double f[3];
        flux_ac(nx, ny, hl, ul, vl, hr, urr, vrr, f);
        if (set_wall) f[0] = ZERO;
        p0 = p0 - f[0] * L; p1 = p1 - f[1] * L; p2 = p2 - f[2] * L;
      } else if (kind == SK_OUT_CRIT) {
        // crit_outflow.f90:32-105
        Face fc = face(g, S, h, i, j);
        double hl = fc.h, ue, ve;
        if (dry_i || hl < h_dry) { ue = ZERO; ve = ZERO; hl = dmax(hl, ZERO); }
        else { ue = fc.u; ve = fc.v; }
        if (hl < h_dry) continue;
        const double tx = g.etx[j * N + i], ty = g.ety[j * N + i];
        const double ul = ue * nx + ve * ny, vl = ue * tx + ve * ty;
        double urr = ul, vrr = vl;
        double hr = urr * urr / G;
        const double a = urr, b = vrr;
        urr = a * nx + b * tx;
        vrr = a * ny + b * ty;
        if (o.hvalve) {
          const double dot = urr * nx + vrr * ny;
          if (dot < ZERO) { hr = hl; urr = ZERO; vrr = ZERO; }
        }
        const double f0 = hr * (urr * nx + vrr * ny);
        const double f1 = hr * urr * (urr * nx + vrr * ny) + HALF * G * hr * hr * nx;
        const double f2 = hr * vrr * (urr * nx + vrr * ny) + HALF * G * hr * hr * ny;
        const double dh = ZERO;
        p0 = p0 - f0 * L; p1 = p1 - (f1 + dh * nx) * L; p2 = p2 - (f2 + dh * ny) * L;
      } else if (kind == SK_OUT_FREE) {
        // free_flow.f90:33-90
        Face fc = face(g, S, h, i, j);
        const double hl = dmax(fc.h, ZERO);
        if (hl < h_dry) continue;
        double ue, ve;
        if (hi < h_dry) { ue = ZERO; ve = ZERO; } else { ue = fc.u; ve = fc.v; }
        double urr = ue, vrr = ve;
        bool set_wall = false;
        if (o.hvalve) {
          const double dot = urr * nx + vrr * ny;
          if (dot < 0) { set_wall = true; urr = ZERO; vrr = ZERO; }
        }
        double f[3];
        flux_ac(nx, ny, hl, ue, ve, hl, urr, vrr, f);
        if (set_wall) f[0] = ZERO;
        p0 = p0 - f[0] * L; p1 = p1 - f[1] * L; p2 = p2 - f[2] * L;
      }
    }
    const double ui = S.u[i], vi = S.v[i];  // stage-input velocities (find_dry_cells)
    double umag_in = ZERO;
    if (!dry_i) umag_in = sqrt(ui * ui + vi * vi);
    if (o.friction == 1) {
      // fric_terms.f90:24-98 (explicit Manning), after the edge fluxes
      double r = dry_i ? ZERO : G * g.cd[i] * g.cd[i] / pow(hi, ONE_THIRD);
      r = -r * umag_in;
      p1 = p1 + r * S.up[i] * g.area[i];
      p2 = p2 + r * S.vp[i] * g.area[i];
    }
    if (!dry_i) {
      // bed_slope_STD.f90:23-70 with adjustZb.f90:41-70
      const double dhx = S.grad[i], dhy = S.grad[N + i];
      double hp[3];
      bool adjust = false;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        hp[k] = hi + dhx * g.vdx[k * N + i] + dhy * g.vdy[k * N + i];
        if (hp[k] < ZERO) adjust = true;
      }
      double gx, gy;
      if (adjust) {
        double zp[3], xx[3], yy[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int v = g.cvtx[k * N + i];
          hp[k] = dmin(ZERO, hp[k]);
          zp[k] = g.zv3[k * N + i] + hp[k];
          xx[k] = g.xv[v]; yy[k] = g.yv[v];
        }
        const double b0 = yy[1] - yy[2], c0 = xx[2] - xx[1];
        const double b1 = yy[2] - yy[0], c1 = xx[0] - xx[2];
        const double b2 = yy[0] - yy[1], c2 = xx[1] - xx[0];
        const double a = HALF / g.area[i];
        gx = a * (b0 * zp[0] + b1 * zp[1] + b2 * zp[2]);
        gy = a * (c0 * zp[0] + c1 * zp[1] + c2 * zp[2]);
      } else {
        gx = g.delzx[i];
        gy = g.delzy[i];
      }
      p1 = p1 - G * hi * gx * g.area[i];
      p2 = p2 - G * hi * gy * g.area[i];
    }
    // solverFVT.f90:259-283
    double delt = c->dt;
    if (o.steady) {
      const double wc = dry_i ? ZERO : sqrt(G * hi);
      double d = umag_in + wc;
      if (d < SMALL) d = 1.044;
      delt = o.cfl * g.area[i] / (d * g.perim[i]);
    }
    const double dt = delt / g.area[i];
    double qh = ZERO, qu = ZERO, qv = ZERO;
    for (int jj = 0; jj < st.s; ++jj) {
      const double a = st.alpha[jj];
      if (a == ZERO) continue;
      const double* sl = st.slot[jj];
      qh = qh + a * sl[i];
      qu = qu + a * sl[N + i];
      qv = qv + a * sl[2 * N + i];
    }
    const double h0n = st.slot[0][i];  // h at the start of the step (residual)
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
      S.u[i] = un;
      S.v[i] = vn;
      double um;
      S.dry_out[i] = find_dry(hn, un, vn, h_dry, h_wet, &um);
    } else {
      // find_dry_cells + fric_terms_impl (solverFVT.f90:335-336)
      double um;
      const bool dpost = find_dry(hn, un, vn, h_dry, h_wet, &um);
      S.dry_post[i] = dpost;
      double rough = ZERO;
      if (!dpost) {
        const double fdt = o.steady ? c->delt_last : c->dt;
        rough = G * g.cd[i] * g.cd[i] / pow(hn, ONE_THIRD);
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
      // hvalve (solverFVT.f90:357-370)
      if (o.hvalve) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const uint32_t inf = g.cinfo[k * N + i];
          if (!(inf & CI_HBAR)) continue;
          const double dot = un * g.enx[k * N + i] + vn * g.eny[k * N + i];
          if (dot < 0) { un = ZERO; vn = ZERO; }
        }
      }
      st.out[i] = qh;
      st.out[N + i] = qu;
      st.out[2 * N + i] = qv;
      S.u[i] = un;
      S.v[i] = vn;
      // phi reset at dry cells (solverFVT.f90:376-386)
      if (dpost) { p0 = ZERO; p1 = ZERO; p2 = ZERO; }
      S.phi[i] = p0; S.phi[N + i] = p1; S.phi[2 * N + i] = p2;
      mxphi = dmax(dmax(dmax(mxphi, fabs(p0)), fabs(p1)), fabs(p2));
      // stage-1 dry flags and CFL candidate of the next step
      double um2;
      const bool dn = find_dry(hn, un, vn, h_dry, h_wet, &um2);
      S.dry_out[i] = dn;
      const double wc = dn ? ZERO : sqrt(G * hn);
      cand = o.cfl * g.area[i] / ((um2 + wc + SMALL) * g.perim[i]);
    }
  }
  if (!st.last || done) return;
  // block partials: deterministic tree over the block (fixed shape)
  __shared__ double s_top[UPD_THREADS], s_bot[UPD_THREADS], s_mx[UPD_THREADS];
  const int t = threadIdx.x;
  s_top[t] = top; s_bot[t] = bot; s_mx[t] = mxphi;
  __syncthreads();
  for (int w = UPD_THREADS / 2; w > 0; w >>= 1) {
    if (t < w) {
      s_top[t] = s_top[t] + s_top[t + w];
      s_bot[t] = s_bot[t] + s_bot[t + w];
      s_mx[t] = dmax(s_mx[t], s_mx[t + w]);
    }
    __syncthreads();
  }
  if (t == 0) {
    part[4 * blockIdx.x] = s_top[0];
    part[4 * blockIdx.x + 1] = s_bot[0];
    part[4 * blockIdx.x + 2] = s_mx[0];
  }
  __syncthreads();
  s_top[t] = cand;
  __syncthreads();
  for (int w = UPD_THREADS / 2; w > 0; w >>= 1) {
    if (t < w) s_top[t] = dmin(s_top[t], s_top[t + w]);
    __syncthreads();
  }
  if (t == 0) part[4 * blockIdx.x + 3] = s_top[0];
}

// to_vertices' cell pass alone (solverFVT.f90:483 -> to_vertices.f90:44-55),
// after k_vertex, for the final state
__global__ void k_final_cells(Geo g, const uint8_t* dry, const uint8_t* vwet, double* u, double* v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.N) return;
  if (dry[i]) return;
  const int N = g.N;
  const bool flag = vwet[g.cvtx[i]] && vwet[g.cvtx[N + i]] && vwet[g.cvtx[2 * N + i]];
  if (!flag) { u[i] = ZERO; v[i] = ZERO; }
}

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
  uint8_t *dry_a = nullptr, *dry_b = nullptr;
  int* d_out_type = nullptr;
  int *d_qts_ptr = nullptr, *d_hts_ptr = nullptr;
  double *d_qts_t = nullptr, *d_qts_v = nullptr, *d_hts_t = nullptr, *d_hts_v = nullptr;
  int nblk_upd = 0;
  int block = 256;
  int iter_host = 0;
  bool started = false;
  bool finalized = false;
  double last_run_s = 0.0;
  // RK order 1 keeps two physical slots and swaps them every step
  int r1_flip = 0;
  // h input of the last stage run: to_vertices after the loop works on the
  // water surface formed at the start of that stage (solverFVT.f90:169,483)
  const double* last_hin = nullptr;

  ~storm_fvt() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* p : owned) (void)hipFree(p);
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

void upload_tables(storm_fvt* C) {
  Prep& P = C->prep;
  auto& own = C->owned;
  C->d_out_type = dupload(own, P.out_type);
  C->d_qts_ptr = dupload(own, P.qts_ptr);
  C->d_hts_ptr = dupload(own, P.hts_ptr);
  C->d_qts_t = dupload(own, P.qts_t);
  C->d_qts_v = dupload(own, P.qts_v);
  C->d_hts_t = dupload(own, P.hts_t);
  C->d_hts_v = dupload(own, P.hts_v);
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
  g.ep1 = dupload(own, P.ep1); g.ep2 = dupload(own, P.ep2);
  g.cinfo = dupload(own, P.cinfo); g.corder = dupload(own, P.corder);
  g.rc = dupload(own, P.rc);
  g.enx = dupload(own, P.enx); g.eny = dupload(own, P.eny);
  g.etx = dupload(own, P.etx); g.ety = dupload(own, P.ety);
  g.elen = dupload(own, P.elen); g.ezmid = dupload(own, P.ezmid);
  g.area = dupload(own, P.area); g.perim = dupload(own, P.perim);
  g.z = dupload(own, P.z); g.cd = dupload(own, P.cd);
  g.delzx = dupload(own, P.delzx); g.delzy = dupload(own, P.delzy);
  g.zhigh = dupload(own, P.zhigh);
  g.vdx = dupload(own, P.vdx); g.vdy = dupload(own, P.vdy); g.zv3 = dupload(own, P.zv3);
  g.mptr = dupload(own, P.mptr); g.mcell = dupload(own, P.mcell);
  g.mdx = dupload(own, P.mdx); g.mdy = dupload(own, P.mdy); g.mw = dupload(own, P.mw);
  g.ai = dupload(own, P.ai); g.bi = dupload(own, P.bi); g.ci = dupload(own, P.ci);
  g.det = dupload(own, P.det);
  g.gconn = dupload(own, P.gconn);
  g.gcx = dupload(own, P.gcx); g.gcy = dupload(own, P.gcy); g.gwt = dupload(own, P.gwt);
  g.vptr = dupload(own, P.vptr); g.vcell = dupload(own, P.vcell);
  g.vbc = dupload(own, P.vbc); g.vw = dupload(own, P.vw); g.zvtx = dupload(own, P.zvtx);
  g.qs_ptr = dupload(own, P.qs_ptr); g.qs_cell = dupload(own, P.qs_cell);
  g.qs_slot = dupload(own, P.qs_slot);
  g.in_type = dupload(own, P.in_type);
  g.n_in = P.n_in; g.n_out = P.n_out; g.last_cell = P.last_cell; g.bch_max = P.bch_max;
  // node coordinates in internal vertex order (adjustZb's partially dry branch)
  g.xv = dupload(own, P.xv_int);
  g.yv = dupload(own, P.yv_int);
  upload_tables(C);
  Opts& o = C->o;
  const storm_fvt_opts_t& q = P.o;
  o.rkorder = q.rkorder; o.dt_op = q.dt_op; o.steady = q.steady; o.stop_cr = q.stop_cr;
  o.btype = q.btype; o.hvalve = q.hvalve; o.limiter = q.limiter; o.flux = q.flux;
  o.gradient = q.gradient; o.friction = q.friction; o.vclip = q.vclip;
  o.cfl = q.cfl; o.rstop = q.rstop; o.h_dry = q.h_dry; o.h_wet = q.h_wet; o.vceiling = q.vceiling;
  const size_t N = P.N;
  const int R = P.R;
  const int nslots = (R == 1) ? 2 : R + 1;
  for (int s = 0; s < nslots && s < 4; ++s)
    if (R == 1 || s >= 1) C->slots[s] = dalloc<double>(own, 3 * N);
  C->block = q.block > 0 ? q.block : 256;
  if (C->block != 64 && C->block != 128 && C->block != 256)
    throw StormErr(STORM_E_ARG, "block must be 64, 128 or 256");
  C->nblk_upd = (int)((N + UPD_THREADS - 1) / UPD_THREADS);
  State& S = C->S;
  S.u = dalloc<double>(own, N); S.v = dalloc<double>(own, N);
  S.up = dalloc<double>(own, N); S.vp = dalloc<double>(own, N);
  S.grad = dalloc<double>(own, 6 * N);
  S.gwork = (q.gradient == STORM_GRAD_GAUSS) ? dalloc<double>(own, 6 * N) : nullptr;
  S.hvtx = dalloc<double>(own, P.P);
  S.zetavtx = dalloc<double>(own, P.P);
  S.vwet = dalloc<uint8_t>(own, P.P);
  S.phi = dalloc<double>(own, 3 * N);
  S.dry_post = dalloc<uint8_t>(own, N);
  S.part = dalloc<double>(own, 4 * (size_t)C->nblk_upd);
  C->dry_a = dalloc<uint8_t>(own, N);
  C->dry_b = dalloc<uint8_t>(own, N);
  S.ctrl = dalloc<Ctrl>(own, 1);
  C->d_h0 = dupload(own, P.h0); C->d_u0 = dupload(own, P.u0); C->d_v0 = dupload(own, P.v0);
}

void start_run(storm_fvt* C) {
  // solverFVT.f90:53-127
  Prep& P = C->prep;
  Ctrl hc;
  std::memset(&hc, 0, sizeof(hc));
  host_time_bcs(P, 0.0, hc.qin, hc.hbc);
  HIPCHK(hipMemcpyAsync(C->S.ctrl, &hc, sizeof(Ctrl), hipMemcpyHostToDevice, C->stream));
  C->r1_flip = 0;
  C->last_hin = C->slot_ptr(0);
  const int nb = (P.N + 255) / 256;
  hipLaunchKernelGGL(k_init, dim3(nb), dim3(256), 0, C->stream, C->g, C->o, C->slot_ptr(0),
                     C->d_h0, C->d_u0, C->d_v0, C->S.u, C->S.v, C->dry_a, C->S.part);
  HIPCHK(hipGetLastError());
  C->S.dry_in = C->dry_a;
  C->S.dry_out = C->dry_b;
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

void launch_control(storm_fvt* C, int begin) {
  hipLaunchKernelGGL(k_control, dim3(1), dim3(CTRL_THREADS), 0, C->stream, C->g, C->o, C->S.ctrl,
                     C->S.part, C->nblk_upd, begin, C->d_out_type, C->d_qts_ptr, C->d_qts_t,
                     C->d_qts_v, C->d_hts_ptr, C->d_hts_t, C->d_hts_v);
}

void launch_step(storm_fvt* C) {
  Prep& P = C->prep;
  const int R = P.R;
  const int N = P.N;
  double a[3][3], b[3];
  rk_table(R, a, b);
  launch_control(C, 1);
  const int bs = C->block;
  const int nbc = (N + bs - 1) / bs;
  const int nbv = (P.P + bs - 1) / bs;
  for (int s = 1; s <= R; ++s) {
    Stage st{};
    for (int j = 0; j < s; ++j) { st.slot[j] = C->slot_ptr(j); st.alpha[j] = a[s - 1][j]; }
    st.out = C->slot_ptr(s);
    st.hin = C->slot_ptr(s - 1);
    st.beta = b[s - 1];
    st.s = s;
    st.last = (s == R);
    hipLaunchKernelGGL(k_vertex, dim3(nbv), dim3(bs), 0, C->stream, C->g, C->o, st.hin,
                       C->S.dry_in, C->S.hvtx, C->S.vwet, (double*)nullptr, C->S.ctrl);
    if (C->o.gradient == STORM_GRAD_GAUSS && C->o.limiter != 0)
      hipLaunchKernelGGL(k_gauss1, dim3(nbc), dim3(bs), 0, C->stream, C->g, st.hin, C->S.u,
                         C->S.v, C->S.gwork, C->S.ctrl);
    hipLaunchKernelGGL(k_grad, dim3(nbc), dim3(bs), 0, C->stream, C->g, C->o, st, C->S);
    if (P.qs_cell.size() > 0)
      hipLaunchKernelGGL(k_bcq, dim3(P.n_in), dim3(BCQ_THREADS), 0, C->stream, C->g, C->o, st, C->S);
    hipLaunchKernelGGL(k_update, dim3(C->nblk_upd), dim3(UPD_THREADS), 0, C->stream, C->g, C->o,
                       st, C->S, C->S.part);
    std::swap(C->S.dry_in, C->S.dry_out);
    C->last_hin = st.hin;
  }
  if (R == 1) C->r1_flip ^= 1;
}

int set_err(storm_fvt* C, int code, const std::string& msg) {
  if (C) C->err = msg;
  return code;
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
}

int storm_fvt_create(const storm_mesh_t* mesh, const storm_fvt_opts_t* opts, storm_fvt_t** ctx) {
  if (!ctx) return STORM_E_ARG;
  *ctx = nullptr;
  std::unique_ptr<storm_fvt> C(new (std::nothrow) storm_fvt());
  if (!C) return STORM_E_NOMEM;
  try {
    int rc = build_prep(mesh, opts, C->prep);
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
  } catch (const StormErr& e) {
    return set_err(C, e.code, e.msg);
  }
  return STORM_OK;
}

int storm_fvt_set_bc_values(storm_fvt_t* C, const double* qin, const double* hbc) {
  if (!C) return STORM_E_ARG;
  try {
    Prep& P = C->prep;
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
    upload_tables(C);
    C->started = false;
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
    HIPCHK(hipEventRecord(C->ev0, C->stream));
    for (int k = 0; k < n; ++k) launch_step(C);
    if (n > 0) launch_control(C, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(C->ev1, C->stream));
    C->iter_host += n;
    Ctrl hc;
    HIPCHK(hipMemcpyAsync(&hc, C->S.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, C->ev0, C->ev1));
    C->last_run_s = ms * 1e-3;
    C->iter_host = hc.iter;
    if (hc.done) C->iter_host = C->prep.o.istop;  // converged: no further steps
    if (iter) *iter = hc.iter;
    if (ttime) *ttime = hc.ttime;
    if (rmax) *rmax = hc.rmax;
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
    // to_vertices after the loop (solverFVT.f90:483): the water surface of
    // the last stage's start and the dry flags of the post-step
    // find_dry_cells (:335)
    const uint8_t* dry = (C->iter_host > 0) ? C->S.dry_post : C->S.dry_in;
    hipLaunchKernelGGL(k_vertex, dim3((P.P + bs - 1) / bs), dim3(bs), 0, C->stream, C->g, C->o,
                       C->last_hin, dry, C->S.hvtx, C->S.vwet, C->S.zetavtx, C->S.ctrl);
    hipLaunchKernelGGL(k_final_cells, dim3((P.N + bs - 1) / bs), dim3(bs), 0, C->stream, C->g,
                       dry, C->S.vwet, C->S.u, C->S.v);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(C->stream));
    C->finalized = true;
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
    if (u) copy_cells(C, C->S.u, u);
    if (v) copy_cells(C, C->S.v, v);
    if (phi)
      for (int c = 0; c < 3; ++c) copy_cells(C, C->S.phi + c * N, phi + c * N);
    if (drycell) {
      // the flags of the post-step find_dry_cells (solverFVT.f90:335)
      std::vector<uint8_t> tmp(N);
      HIPCHK(hipMemcpy(tmp.data(), (C->iter_host > 0) ? C->S.dry_post : C->S.dry_in, N,
                       hipMemcpyDeviceToHost));
      for (size_t ii = 0; ii < N; ++ii) drycell[P.perm[ii]] = tmp[ii];
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

int storm_fvt_get_vertex(storm_fvt_t* C, double* hvtx, double* zetavtx) {
  if (!C) return STORM_E_ARG;
  try {
    HIPCHK(hipStreamSynchronize(C->stream));
    Prep& P = C->prep;
    std::vector<double> tmp(P.P);
    if (hvtx) {
      HIPCHK(hipMemcpy(tmp.data(), C->S.hvtx, sizeof(double) * P.P, hipMemcpyDeviceToHost));
      for (int pp = 0; pp < P.P; ++pp) hvtx[P.vperm[pp]] = tmp[pp];
    }
    if (zetavtx) {
      HIPCHK(hipMemcpy(tmp.data(), C->S.zetavtx, sizeof(double) * P.P, hipMemcpyDeviceToHost));
      for (int pp = 0; pp < P.P; ++pp) zetavtx[P.vperm[pp]] = tmp[pp];
    }
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
