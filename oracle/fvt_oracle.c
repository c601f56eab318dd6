/*
 * TEST INFRASTRUCTURE ONLY -- CPU oracle of the SToRM FVT time step.
 *
 * Plain-C restatement of the reference algorithm (Source_xsy-64bit/, cited
 * per function as file:line).  It keeps the reference's loop orders and
 * expression orders so that, built without FMA contraction, it reproduces the
 * flang-built reference bit for bit; the tests check that.  Used only as the
 * checker of the HIP path and as the timed CPU baseline; the product never
 * links it.
 *
 * Parity status per function (what pins it):
 *  - set-up, dataprepFVT's module copies, find_dry_cells, lsq_grad_array,
 *    limiterBJ, hgradfix2, to_vertices (vertex velocities and enforce_bcs
 *    included), invisc_fluxes+flux_ac, wenslp3/weslp2, inflowQbyH, inflowH,
 *    outflowbyH, crit_outflow, free_flow, bed_slope_STD/adjustZb, RK update
 *    and wet/dry recovery, fric_terms_impl, residual, hvalve, ctr2vtx of cd,
 *    error_est(2): bit-for-bit against the flang-built reference on the
 *    committed whole-run cases.
 *  - flux_rusanov, fric_terms (explicit), fvsources, part_dry, drynwet:
 *    bit-for-bit against the reference routines themselves, run by the
 *    per-routine harness ref_units.f90 on seeded inputs.
 *  - flux_hll (textbook form; flux_HLL.f90's star region and both-wet right
 *    speed are defective): its dry-bed speeds and upwind fluxes bit-for-bit
 *    against the reference routine; the star region by closed form.
 *  - flux_bgnvc (corrected eigenvectors), gauss gradients (corrected
 *    gaussgrad_setup.f90; ghost cells by the mirror identity), the corrected
 *    CFL time step: closed forms (consistency, Roe identities, exactness on
 *    linear fields, brute-force minimum).
 */
#include "fvt_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define G 9.81
#define HALF 0.5
#define ONE 1.0
#define ZERO 0.0
#define RHO 998.60  /* constants.f90:29 */
static const double ONE_THIRD = 1.0 / 3.0;

/* Sensitivity probe (test use only, off by default): with probability p per
 * call, the friction's h**(1/3) is moved by one ulp, as another libm's pow
 * may round it.  Drawn from a xorshift64 stream seeded by the caller. */
static double sens_p = 0.0;
static uint64_t sens_state = 1;
void orc_set_pow_perturbation(double p, uint64_t seed) {
  sens_p = p;
  sens_state = seed | 1;
}
static double fric_pow(double h) {
  double r = pow(h, ONE_THIRD);
  if (sens_p > 0.0) {
    sens_state ^= sens_state << 13;
    sens_state ^= sens_state >> 7;
    sens_state ^= sens_state << 17;
    if ((double)(sens_state >> 11) * (1.0 / 9007199254740992.0) < sens_p)
      r = nextafter(r, (sens_state & 1) ? INFINITY : -INFINITY);
  }
  return r;
}
/* REAL(mp), PARAMETER :: small = 1.0E-9 -- a default-real literal
 * (solverFVT.f90:44, deltatFVT.f90:30, flux_ac.f90:46). */
static const double SMALL = (double)1.0e-9f;

struct orc {
  int P, N, E;
  orc_opts_t o;
  /* geometry */
  double *x, *y;
  int *vtx;                /* [3N] */
  int *gedge;              /* [3N] edge id */
  int *gsign;              /* [3N] +1/-1 orientation */
  double *area, *perim, *xc, *yc;
  int *ep, *ee;            /* [2E] */
  double *nrm, *tng, *len; /* [2E],[2E],[E] */
  int nflow1, *flowedg1;
  int nwall1, *walledg1;
  int nbpoly, *bpolygon;
  int nbc, *bcedges;
  unsigned char *hbar, *qbar;
  /* boundary strings */
  int n_in, n_out;
  int *in_type, *in_ptr, *in_edges;   /* edges between consecutive nodes */
  int *in_nodes_ptr, *in_nodes;
  int *out_type, *out_ptr, *out_edges;
  int *out_nodes_ptr, *out_nodes;
  int *qts_ptr, *hts_ptr;
  double *qts_t, *qts_v, *hts_t, *hts_v;
  double *qin, *hbc;
  double bch_max;
  /* enforce_bcs tables: per string node the normal of ibc_arrays.f90:93-135
   * (inflow, inward) and :183-253 (outflow, outward), per inflow string edge
   * its length (l_inflow); wall points of walls.f90:150-176 and their
   * tangents (ibc_arrays.f90:266-330) */
  double *pn, *lin, *hn;
  int n_wall, *wall_pts;
  double *wtang;
  /* post-loop fields (solverFVT.f90:483-487) */
  double *uvtx, *vvtx, *cdvtx, *e_est;
  /* dataprepFVT.f90:69-140's module copies, kept as set-up left them */
  double *p_zeta, *p_cdvtx, *p_hvtx, *p_uvtx, *p_vvtx, *p_zetavtx;
  /* topology */
  int *n2t_ptr, *n2t;
  int *t2t_ptr, *t2t;
  int *t2t3;               /* [3N] */
  /* dataprep */
  double *rc;              /* [N][3][2] */
  double *c2v_w, *c2v_wid; /* aligned with n2t */
  double *lsq_dx, *lsq_dy, *lsq_w; /* aligned with t2t */
  double *ai, *bi, *ci, *det;
  int *csortedz;           /* [3N] */
  int *gconn;              /* [3N] gauss */
  double *gcx, *gcy, *gwt; /* [3N] gauss */
  double rka[3][3], rkb[3];
  /* elemnt_db.f90:42-128's t2t kept for the residual smoothing, which reads
   * it whatever the gradients use (solverFVT.f90:249) */
  int *et2t_ptr, *et2t;
  /* wind (windforcing.f90:64-86): per cell, NULL when off */
  double *wcoef1, *wcoef2;
  /* culverts: cvprep.f90's cells and offsets, ioculvert.f90's curves */
  int ncv, *cvin, *cvout, *cvtab_ptr;
  double *cvpin, *cvrin, *cvhead, *cvq, *cvsrc;
  double *phi0;            /* [3N] unsmoothed residual */
  /* state */
  double *z, *zvtx, *cd;
  double *h, *u, *v, *zeta, *u_mag, *wavec, *rough;
  double *hvtx, *zetavtx;
  double *delh, *delu, *delv, *delz, *gwork; /* [2N] */
  double *phi;             /* [3N] component-major: phi[c*N+i] */
  unsigned char *dry;
  double *rk;              /* [3][rkorder+1][N]: rk[(var*(R+1)+slot)*N+i] */
  int iter;
  double ttime, stime, rmax, delt;
  int started;
};

#define RK(s, var, slot, i) ((s)->rk[((size_t)(var) * ((s)->o.rkorder + 1) + (slot)) * (size_t)(s)->N + (i)])

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

/* equals.f90:14-24 */
int32_t orc_equals(double a, double b) {
  double eps = fabs(a) * DBL_EPSILON;
  if (eps == ZERO) eps = DBL_MIN;
  return (fabs(a - b) > eps) ? 0 : 1;
}

static double fmin2(double a, double b) { return (a < b) ? a : b; }
static double fmax2(double a, double b) { return (a > b) ? a : b; }

/* local_edge.f90:32-39 */
static int local_edge(const orc_t *s, int cell, int k) {
  for (int j = 0; j < 3; ++j)
    if (s->gedge[3 * cell + j] == k) return j;
  return -1;
}

/* edge_in_element.f90:25-26 */
static int edge_in_element(const orc_t *s, int k) {
  int i = s->ee[2 * k];
  if (i < 0) i = s->ee[2 * k + 1];
  return i;
}

/* ======================================================================
 * Set-up
 * ====================================================================== */

/* cde_t2t.f90:30-47: remove duplicate entries of row i in place.  As in the
 * Fortran DO loop, the trip count of the inner sweep is fixed when the sweep
 * starts, even though n shrinks inside it. */
static void cde_t2t(int *row, int *n_io) {
  int n = *n_io;
  int l = 1;
  for (;;) {
    int j = row[l];
    const int mend = n;
    for (int m = l + 1; m <= mend; ++m) {
      int k = row[m];
      if (j == k) {
        n = n - 1;
        for (int nn = m; nn <= n; ++nn) row[nn] = row[nn + 1];
      }
    }
    l = l + 1;
    if (l > n) break;
  }
  *n_io = n;
}

/* elemnt_db.f90:136-214: the molecule of the higher-order least squares
 * (ACTIVATE HDLS), which replaces t2t for the gradients.  Row i lists every
 * other triangle on i's three vertices, vertex by vertex in n2t order, then
 * cde_t2t drops repeats.  A row left with fewer than three entries is
 * augmented as :168-197 write it: the loop index j = 2..nn is used as a
 * triangle NUMBER (grid(j)), and a candidate is skipped when its column in
 * the n2t row equals j. */
static int build_t2thd(orc_t *s, char *err, int errlen) {
  const int N = s->N, P = s->P;
  int maxc = 0;
  for (int p = 0; p < P; ++p) {
    int c = s->n2t_ptr[p + 1] - s->n2t_ptr[p];
    if (c > maxc) maxc = c;
  }
  const int MAXD = 4 * maxc + 3; /* columns, the count included */
  int *w = xcalloc((size_t)N * MAXD, sizeof(int));
#define W(i, k) w[(size_t)(i) * MAXD + (k)]
  for (int i = 0; i < N; ++i) {
    for (int j = 0; j < 3; ++j) {
      int k = s->vtx[3 * i + j];
      for (int t = s->n2t_ptr[k]; t < s->n2t_ptr[k + 1]; ++t) {
        int m = s->n2t[t];
        if (m == i) continue;
        W(i, W(i, 0) + 1) = m;
        W(i, 0) = W(i, 0) + 1;
      }
    }
    int cnt = W(i, 0);
    cde_t2t(&W(i, 0), &cnt);
    W(i, 0) = cnt;
  }
  for (int i = 0; i < N; ++i) {
    int n = W(i, 0);
    if (n > 2) continue;
    if (n < 1) { free(w); snprintf(err, errlen, "triangle connectivity table incorrect"); return -1; }
    int nn = W(i, 0) + 1 < 3 ? W(i, 0) + 1 : 3;
    for (int j = 2; j <= nn; ++j) {     /* j: a 1-based triangle number */
      for (int k = 0; k < 3; ++k) {
        int l = s->vtx[3 * (j - 1) + k];
        int cl = s->n2t_ptr[l + 1] - s->n2t_ptr[l];
        for (int m = 2; m <= cl + 1; ++m) { /* m: 1-based column of the n2t row */
          int t = s->n2t[s->n2t_ptr[l] + m - 2];
          if (t == i || m == j) continue;
          if (W(i, 0) + 2 > MAXD) { free(w); snprintf(err, errlen, "allocation bounds in T2TW"); return -1; }
          W(i, W(i, 0) + 1) = t;
          W(i, 0) = W(i, 0) + 1;
        }
      }
    }
    int cnt = W(i, 0);
    cde_t2t(&W(i, 0), &cnt);
    W(i, 0) = cnt;
  }
  free(s->t2t_ptr);
  free(s->t2t);
  s->t2t_ptr = xcalloc(N + 1, sizeof(int));
  for (int i = 0; i < N; ++i) s->t2t_ptr[i + 1] = s->t2t_ptr[i] + W(i, 0);
  s->t2t = xcalloc(s->t2t_ptr[N] + 1, sizeof(int));
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < W(i, 0); ++k) s->t2t[s->t2t_ptr[i] + k] = W(i, k + 1);
#undef W
  free(w);
  return 0;
}

static int build(orc_t *s, const orc_case_t *c, char *err, int errlen) {
  const int P = s->P, N = s->N, E = s->E;
  /* nodes and connectivity (tecplot_read.f90:50-143) */
  memcpy(s->x, c->x, sizeof(double) * P);
  memcpy(s->y, c->y, sizeof(double) * P);
  memcpy(s->vtx, c->tri, sizeof(int) * 3 * N);
  for (int k = 0; k < E; ++k) {
    s->ep[2 * k] = c->edge_p[2 * k];
    s->ep[2 * k + 1] = c->edge_p[2 * k + 1];
    int e1 = c->edge_e[2 * k], e2 = c->edge_e[2 * k + 1];
    if (e1 < 0) { e1 = e2; e2 = -1; }  /* tecplot_read.f90:139-142 */
    s->ee[2 * k] = e1;
    s->ee[2 * k + 1] = e2;
  }
  /* find_edges.f90:74-114 (edge data given): signed local edges */
  for (int i = 0; i < 3 * N; ++i) { s->gedge[i] = -1; s->gsign[i] = 0; }
  for (int j = 0; j < E; ++j) {
    for (int kk = 0; kk < 2; ++kk) {
      int i = s->ee[2 * j + kk];
      if (i < 0) continue;
      for (int l = 0; l < 3; ++l) {
        int a = s->vtx[3 * i + l], b = s->vtx[3 * i + (l + 1) % 3];
        if (a == s->ep[2 * j] && b == s->ep[2 * j + 1]) {
          s->gedge[3 * i + l] = j; s->gsign[3 * i + l] = 1; break;
        } else if (a == s->ep[2 * j + 1] && b == s->ep[2 * j]) {
          s->gedge[3 * i + l] = j; s->gsign[3 * i + l] = -1; break;
        }
      }
    }
  }
  for (int i = 0; i < 3 * N; ++i)
    if (s->gedge[i] < 0) { snprintf(err, errlen, "edge data structure incomplete"); return -1; }
  /* edge_normals.f90:30-39, find_edges.f90:159-163 */
  for (int k = 0; k < E; ++k) {
    double x1 = s->x[s->ep[2 * k]], y1 = s->y[s->ep[2 * k]];
    double x2 = s->x[s->ep[2 * k + 1]], y2 = s->y[s->ep[2 * k + 1]];
    s->nrm[2 * k] = y2 - y1;
    s->nrm[2 * k + 1] = x1 - x2;
    s->tng[2 * k] = x2 - x1;
    s->tng[2 * k + 1] = y2 - y1;
    double dx = x1 - x2, dy = y1 - y2;
    s->len[k] = sqrt(dx * dx + dy * dy);
  }
  /* t_areas.f90:33-52 */
  for (int i = 0; i < N; ++i) {
    int a = s->vtx[3 * i], b = s->vtx[3 * i + 1], d = s->vtx[3 * i + 2];
    double x1 = s->x[a], y1 = s->y[a], x2 = s->x[b], y2 = s->y[b], x3 = s->x[d], y3 = s->y[d];
    s->area[i] = HALF * fabs((x1 - x2) * (y1 + y2) + (x2 - x3) * (y2 + y3) + (x3 - x1) * (y3 + y1));
    s->xc[i] = ONE_THIRD * (x1 + x2 + x3);
    s->yc[i] = ONE_THIRD * (y1 + y2 + y3);
    s->perim[i] = s->len[s->gedge[3 * i]] + s->len[s->gedge[3 * i + 1]] + s->len[s->gedge[3 * i + 2]];
  }
  /* boundary strings (ioflow.f90) */
  s->n_in = c->n_in;
  s->n_out = c->n_out;
  unsigned char *isbc = xcalloc(P, 1);
  int nin_nodes = c->n_in ? c->in_ptr[c->n_in] : 0;
  int nout_nodes = c->n_out ? c->out_ptr[c->n_out] : 0;
  for (int t = 0; t < nin_nodes; ++t) isbc[c->in_nodes[t]] = 1;
  for (int t = 0; t < nout_nodes; ++t) isbc[c->out_nodes[t]] = 1;
  /* walls.f90:40-148 */
  int *cnt = xcalloc(E, sizeof(int));
  for (int n = 0; n < N; ++n)
    for (int j = 0; j < 3; ++j) cnt[s->gedge[3 * n + j]]++;
  s->flowedg1 = xcalloc(E, sizeof(int));
  s->walledg1 = xcalloc(E, sizeof(int));
  s->bpolygon = xcalloc(E, sizeof(int));
  s->nflow1 = s->nwall1 = s->nbpoly = 0;
  for (int i = 0; i < E; ++i) {
    if (cnt[i] == 1) {
      int p1 = s->ep[2 * i], p2 = s->ep[2 * i + 1];
      s->bpolygon[s->nbpoly++] = i;
      if (!(isbc[p1] && isbc[p2])) s->walledg1[s->nwall1++] = i;
    } else if (cnt[i] == 2) {
      s->flowedg1[s->nflow1++] = i;
    } else {
      snprintf(err, errlen, "grid construction error in edge assignment");
      free(cnt); free(isbc); return -1;
    }
  }
  free(cnt);
  /* ibc_arrays.f90:59-181: edge chains of the boundary strings (is_edge) */
  s->in_type = xcalloc(s->n_in + 1, sizeof(int));
  s->in_ptr = xcalloc(s->n_in + 1, sizeof(int));
  s->in_edges = xcalloc(nin_nodes + 1, sizeof(int));
  s->in_nodes_ptr = xcalloc(s->n_in + 1, sizeof(int));
  s->in_nodes = xcalloc(nin_nodes + 1, sizeof(int));
  s->out_type = xcalloc(s->n_out + 1, sizeof(int));
  s->out_ptr = xcalloc(s->n_out + 1, sizeof(int));
  s->out_edges = xcalloc(nout_nodes + 1, sizeof(int));
  s->out_nodes_ptr = xcalloc(s->n_out + 1, sizeof(int));
  s->out_nodes = xcalloc(nout_nodes + 1, sizeof(int));
  /* node -> edge lookup via n2t is built below; use a direct search here */
  for (int ch = 0, ne = 0; ch < s->n_in; ++ch) {
    s->in_type[ch] = c->in_type[ch];
    s->in_ptr[ch] = ne;
    s->in_nodes_ptr[ch] = c->in_ptr[ch];
    for (int t = c->in_ptr[ch]; t < c->in_ptr[ch + 1]; ++t) s->in_nodes[t] = c->in_nodes[t];
    for (int t = c->in_ptr[ch]; t + 1 < c->in_ptr[ch + 1]; ++t) {
      int p1 = c->in_nodes[t], p2 = c->in_nodes[t + 1], e = -1;
      for (int k = 0; k < E; ++k)
        if ((s->ep[2 * k] == p1 && s->ep[2 * k + 1] == p2) || (s->ep[2 * k] == p2 && s->ep[2 * k + 1] == p1)) { e = k; break; }
      if (e < 0) { snprintf(err, errlen, "inflow nodes %d %d not connected by an edge", p1, p2); free(isbc); return -1; }
      s->in_edges[ne++] = e;
    }
    s->in_ptr[ch + 1] = ne;
  }
  s->in_nodes_ptr[s->n_in] = nin_nodes;
  for (int ch = 0, ne = 0; ch < s->n_out; ++ch) {
    s->out_type[ch] = c->out_type[ch];
    s->out_ptr[ch] = ne;
    s->out_nodes_ptr[ch] = c->out_ptr[ch];
    for (int t = c->out_ptr[ch]; t < c->out_ptr[ch + 1]; ++t) s->out_nodes[t] = c->out_nodes[t];
    for (int t = c->out_ptr[ch]; t + 1 < c->out_ptr[ch + 1]; ++t) {
      int p1 = c->out_nodes[t], p2 = c->out_nodes[t + 1], e = -1;
      for (int k = 0; k < E; ++k)
        if ((s->ep[2 * k] == p1 && s->ep[2 * k + 1] == p2) || (s->ep[2 * k] == p2 && s->ep[2 * k + 1] == p1)) { e = k; break; }
      if (e < 0) { snprintf(err, errlen, "outflow nodes %d %d not connected by an edge", p1, p2); free(isbc); return -1; }
      s->out_edges[ne++] = e;
    }
    s->out_ptr[ch + 1] = ne;
  }
  s->out_nodes_ptr[s->n_out] = nout_nodes;
  free(isbc);
  /* ibc_arrays.f90:377-466: bcedges, qbar, hbar */
  s->bcedges = xcalloc(s->nbpoly + 1, sizeof(int));
  s->nbc = 0;
  {
    unsigned char *iswall = xcalloc(E, 1);
    for (int t = 0; t < s->nwall1; ++t) iswall[s->walledg1[t]] = 1;
    for (int t = 0; t < s->nbpoly; ++t)
      if (!iswall[s->bpolygon[t]]) s->bcedges[s->nbc++] = s->bpolygon[t];
    free(iswall);
  }
  s->hbar = xcalloc(s->nbc + 1, 1);
  s->qbar = xcalloc(s->nbc + 1, 1);
  for (int t = 0; t < (s->n_in ? s->in_ptr[s->n_in] : 0); ++t)
    for (int j = 0; j < s->nbc; ++j)
      if (s->in_edges[t] == s->bcedges[j]) { s->qbar[j] = 1; break; }
  for (int t = 0; t < (s->n_out ? s->out_ptr[s->n_out] : 0); ++t)
    for (int j = 0; j < s->nbc; ++j)
      if (s->out_edges[t] == s->bcedges[j]) { s->hbar[j] = 1; break; }
  /* time tables */
  s->qts_ptr = xcalloc(s->n_in + 1, sizeof(int));
  s->hts_ptr = xcalloc(s->n_out + 1, sizeof(int));
  int nq = s->n_in ? c->qts_ptr[s->n_in] : 0, nh = s->n_out ? c->hts_ptr[s->n_out] : 0;
  s->qts_t = xcalloc(nq + 1, sizeof(double)); s->qts_v = xcalloc(nq + 1, sizeof(double));
  s->hts_t = xcalloc(nh + 1, sizeof(double)); s->hts_v = xcalloc(nh + 1, sizeof(double));
  if (s->n_in) memcpy(s->qts_ptr, c->qts_ptr, sizeof(int) * (s->n_in + 1));
  if (s->n_out) memcpy(s->hts_ptr, c->hts_ptr, sizeof(int) * (s->n_out + 1));
  if (nq) { memcpy(s->qts_t, c->qts_t, sizeof(double) * nq); memcpy(s->qts_v, c->qts_v, sizeof(double) * nq); }
  if (nh) { memcpy(s->hts_t, c->hts_t, sizeof(double) * nh); memcpy(s->hts_v, c->hts_v, sizeof(double) * nh); }
  s->qin = xcalloc(s->n_in + 1, sizeof(double));
  s->hbc = xcalloc(s->n_out + 1, sizeof(double));
  /* ibc_arrays.f90:93-135: inflow string node normals.  They are formed
   * before dataprepFVT.f90:176-181 makes the edge normals unit, so each
   * edge's normal still carries its length. */
  s->pn = xcalloc(2 * (size_t)nin_nodes + 2, sizeof(double));
  s->lin = xcalloc((size_t)nin_nodes + 1, sizeof(double));
  for (int ch = 0; ch < s->n_in; ++ch) {
    const int a = s->in_nodes_ptr[ch], n = s->in_nodes_ptr[ch + 1] - a;
    double *pn = s->pn + 2 * (size_t)a;
    for (int i = 0; i + 1 < n; ++i) {
      const int e = s->in_edges[s->in_ptr[ch] + i];
      const int t = edge_in_element(s, e);
      const int j = local_edge(s, t, e);
      if (s->gsign[3 * t + j] > 0) { pn[2 * i] = -s->nrm[2 * e]; pn[2 * i + 1] = -s->nrm[2 * e + 1]; }
      else { pn[2 * i] = s->nrm[2 * e]; pn[2 * i + 1] = s->nrm[2 * e + 1]; }
      s->lin[a + i] = s->len[e];
    }
    if (n > 1) { pn[2 * (n - 1)] = pn[2 * (n - 2)]; pn[2 * (n - 1) + 1] = pn[2 * (n - 2) + 1]; }
    for (int i = n - 2; i >= 1; --i) {
      pn[2 * i] = pn[2 * i] + pn[2 * (i - 1)];
      pn[2 * i + 1] = pn[2 * i + 1] + pn[2 * (i - 1) + 1];
    }
    for (int i = 0; i < n; ++i) {
      const double l = sqrt(pn[2 * i] * pn[2 * i] + pn[2 * i + 1] * pn[2 * i + 1]);
      pn[2 * i] = pn[2 * i] / l;
      pn[2 * i + 1] = pn[2 * i + 1] / l;
    }
  }
  /* ibc_arrays.f90:183-253 (outflow string nodes) and :266-330 (wall
   * points): the sum of the outward normals of the first two boundary
   * polygon edges through the node, made unit (for a wall point then turned
   * by +90 degrees into a tangent) */
  s->hn = xcalloc(2 * (size_t)nout_nodes + 2, sizeof(double));
  {
    int *isstr = xcalloc(P, sizeof(int));
    for (int t = 0; t < nin_nodes; ++t) isstr[c->in_nodes[t]] = 1;
    for (int t = 0; t < nout_nodes; ++t) isstr[c->out_nodes[t]] = 1;
    /* walls.f90:150-176 */
    int *seen = xcalloc(P, sizeof(int));
    s->wall_pts = xcalloc(2 * (size_t)s->nwall1 + 1, sizeof(int));
    s->n_wall = 0;
    for (int i = 0; i < s->nwall1; ++i)
      for (int j = 0; j < 2; ++j) {
        const int p = s->ep[2 * s->walledg1[i] + j];
        if (!seen[p] && !isstr[p]) { seen[p] = 1; s->wall_pts[s->n_wall++] = p; }
      }
    free(seen);
    free(isstr);
    s->wtang = xcalloc(2 * (size_t)s->n_wall + 2, sizeof(double));
    for (int q = 0; q < nout_nodes + s->n_wall; ++q) {
      const int node = (q < nout_nodes) ? s->out_nodes[q] : s->wall_pts[q - nout_nodes];
      int e1 = -1, e2 = -1, counter = 0;
      for (int j = 0; j < s->nbpoly; ++j) {
        const int k = s->bpolygon[j];
        if (node == s->ep[2 * k] || node == s->ep[2 * k + 1]) {
          e2 = k;
          counter = counter + 1;
          if (counter == 2) break;
          e1 = k;
        }
      }
      if (counter == 0) { snprintf(err, errlen, "string or wall node %d on no boundary edge", node); return -1; }
      double nn[2][2];
      for (int w = 0; w < 2; ++w) {
        const int e = w ? e2 : e1;
        const int t = edge_in_element(s, e);
        const double sg = (s->gsign[3 * t + local_edge(s, t, e)] > 0) ? ONE : -ONE;
        nn[w][0] = (sg > 0) ? s->nrm[2 * e] : -s->nrm[2 * e];
        nn[w][1] = (sg > 0) ? s->nrm[2 * e + 1] : -s->nrm[2 * e + 1];
      }
      double nx = nn[0][0] + nn[1][0], ny = nn[0][1] + nn[1][1];
      const double l = sqrt(nx * nx + ny * ny);
      nx = nx / l;
      ny = ny / l;
      if (q < nout_nodes) { s->hn[2 * q] = nx; s->hn[2 * q + 1] = ny; }
      else { const int w = q - nout_nodes; s->wtang[2 * w] = -ny; s->wtang[2 * w + 1] = nx; }
    }
  }

  /* node_db.f90:32-68: n2t, cells in ascending order */
  s->n2t_ptr = xcalloc(P + 1, sizeof(int));
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < 3; ++j) s->n2t_ptr[s->vtx[3 * i + j] + 1]++;
  for (int p = 0; p < P; ++p) s->n2t_ptr[p + 1] += s->n2t_ptr[p];
  s->n2t = xcalloc(s->n2t_ptr[P] + 1, sizeof(int));
  {
    int *fill = xcalloc(P, sizeof(int));
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < 3; ++j) {
        int p = s->vtx[3 * i + j];
        s->n2t[s->n2t_ptr[p] + fill[p]++] = i;
      }
    free(fill);
  }
  /* elemnt_db.f90:42-128: t2t, with the boundary molecule augmentation */
  if (N < 4) { snprintf(err, errlen, "too few triangles (< 4)"); return -1; }
  {
    const int MAXD = 20;
    int *w = xcalloc((size_t)N * (MAXD + 1), sizeof(int)); /* row i: w[i*(MAXD+1)+0] = count */
#define W(i, k) w[(size_t)(i) * (MAXD + 1) + (k)]
    for (int i = 0; i < E; ++i) {
      int e1 = s->ee[2 * i], e2 = s->ee[2 * i + 1];
      if (e1 < 0 || e2 < 0) continue;
      int k = W(e1, 0) + 1; W(e1, k) = e2; W(e1, 0) = k;
      k = W(e2, 0) + 1; W(e2, k) = e1; W(e2, 0) = k;
    }
    for (int i = 0; i < N; ++i) {
      int n = W(i, 0);
      if (n == 3) continue;
      if (n < 1 || n > 3) { free(w); snprintf(err, errlen, "triangle connectivity table incorrect"); return -1; }
      for (int l = 1; l <= n; ++l) {
        int j = W(i, l);
        int mm = W(j, 0) < 3 ? W(j, 0) : 3;
        for (int m = 1; m <= mm; ++m) {
          int k = W(j, m);
          if (k == i) continue;
          int nn = W(i, 0) + 1;
          if (nn > MAXD) { free(w); snprintf(err, errlen, "t2t overflow"); return -1; }
          W(i, nn) = k; W(i, 0) = nn;
        }
      }
      int cnt2 = W(i, 0);
      cde_t2t(&W(i, 0), &cnt2);
      W(i, 0) = cnt2;
    }
    for (int i = 0; i < N; ++i) {
      int n = W(i, 0);
      if (n >= 3) continue;
      int j = W(i, n);
      int mm = W(j, 0);
      for (int m = 1; m <= mm; ++m) {
        int k = W(j, m);
        if (k == i) continue;
        int nn = W(i, 0) + 1;
        if (nn > MAXD) { free(w); snprintf(err, errlen, "t2t overflow"); return -1; }
        W(i, nn) = k; W(i, 0) = nn;
      }
      int cnt2 = W(i, 0);
      cde_t2t(&W(i, 0), &cnt2);
      W(i, 0) = cnt2;
    }
    s->t2t_ptr = xcalloc(N + 1, sizeof(int));
    for (int i = 0; i < N; ++i) s->t2t_ptr[i + 1] = s->t2t_ptr[i] + W(i, 0);
    s->t2t = xcalloc(s->t2t_ptr[N] + 1, sizeof(int));
    for (int i = 0; i < N; ++i)
      for (int k = 0; k < W(i, 0); ++k) s->t2t[s->t2t_ptr[i] + k] = W(i, k + 1);
#undef W
    free(w);
  }
  /* elemnt_db.f90:226-233: edge neighbours per local edge */
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < 3; ++j) {
      int k = s->gedge[3 * i + j];
      int e1 = s->ee[2 * k];
      if (e1 == i) e1 = s->ee[2 * k + 1];
      s->t2t3[3 * i + j] = e1;
    }
  s->et2t_ptr = xcalloc(N + 1, sizeof(int));
  memcpy(s->et2t_ptr, s->t2t_ptr, sizeof(int) * (N + 1));
  s->et2t = xcalloc(s->t2t_ptr[N] + 1, sizeof(int));
  memcpy(s->et2t, s->t2t, sizeof(int) * s->t2t_ptr[N]);
  if (s->o.gradient == ORC_GRAD_HDLS) return build_t2thd(s, err, errlen);
  return 0;
}

/* vtx2ctr.f90:32-35 */
static void vtx2ctr(const orc_t *s, const double *va, double *ca) {
  for (int i = 0; i < s->N; ++i)
    ca[i] = (va[s->vtx[3 * i]] + va[s->vtx[3 * i + 1]] + va[s->vtx[3 * i + 2]]) * ONE_THIRD;
}

/* ctr2vtx.f90:73-80 (type 3, normalised inverse-distance weights) */
static void ctr2vtx3(const orc_t *s, const double *ca, double *va) {
  for (int i = 0; i < s->P; ++i) {
    double a = ZERO;
    for (int j = s->n2t_ptr[i]; j < s->n2t_ptr[i + 1]; ++j) a = a + s->c2v_w[j] * ca[s->n2t[j]];
    va[i] = a;
  }
}

static int dataprep(orc_t *s, const orc_case_t *c) {
  const int P = s->P, N = s->N, E = s->E;
  /* c2v_setup.f90:106-119 (type 3) */
  for (int i = 0; i < P; ++i) {
    double tdist = ZERO;
    for (int j = s->n2t_ptr[i]; j < s->n2t_ptr[i + 1]; ++j) {
      int k = s->n2t[j];
      double xk = s->xc[k] - s->x[i], yk = s->yc[k] - s->y[i];
      s->c2v_w[j] = ONE / sqrt(xk * xk + yk * yk);
      s->c2v_wid[j] = s->c2v_w[j];
      tdist = tdist + s->c2v_w[j];
    }
    for (int j = s->n2t_ptr[i]; j < s->n2t_ptr[i + 1]; ++j) s->c2v_w[j] = s->c2v_w[j] / tdist;
  }
  /* dataprepFVT.f90:73-140 */
  double *tmpv = xcalloc(P > N ? P : N, sizeof(double));
  if (c->z_at_vtx) {
    vtx2ctr(s, c->z, s->z);
    memcpy(s->zvtx, c->z, sizeof(double) * P);
  } else {
    memcpy(s->z, c->z, sizeof(double) * N);
    ctr2vtx3(s, c->z, s->zvtx);
    vtx2ctr(s, s->zvtx, s->z);
  }
  if (c->h_at_vtx) {
    vtx2ctr(s, c->h, s->h);
    for (int i = 0; i < P; ++i) { s->hvtx[i] = c->h[i]; s->zetavtx[i] = s->hvtx[i] + s->zvtx[i]; }
    vtx2ctr(s, s->zetavtx, s->zeta);
  } else {
    memcpy(s->h, c->h, sizeof(double) * N);
    for (int i = 0; i < N; ++i) s->zeta[i] = s->z[i] + s->h[i];
    ctr2vtx3(s, s->zeta, s->zetavtx);
    for (int i = 0; i < P; ++i) s->hvtx[i] = s->zetavtx[i] - s->zvtx[i];
  }
  if (c->u_at_vtx) { vtx2ctr(s, c->u, s->u); memcpy(s->p_uvtx, c->u, sizeof(double) * P); }
  else { memcpy(s->u, c->u, sizeof(double) * N); ctr2vtx3(s, c->u, s->p_uvtx); }
  if (c->v_at_vtx) { vtx2ctr(s, c->v, s->v); memcpy(s->p_vvtx, c->v, sizeof(double) * P); }
  else { memcpy(s->v, c->v, sizeof(double) * N); ctr2vtx3(s, c->v, s->p_vvtx); }
  if (c->cd_at_vtx) { vtx2ctr(s, c->cd, s->cd); memcpy(s->p_cdvtx, c->cd, sizeof(double) * P); }
  else { memcpy(s->cd, c->cd, sizeof(double) * N); ctr2vtx3(s, c->cd, s->p_cdvtx); }
  memcpy(s->p_zeta, s->zeta, sizeof(double) * N);
  memcpy(s->p_hvtx, s->hvtx, sizeof(double) * P);
  memcpy(s->p_zetavtx, s->zetavtx, sizeof(double) * P);
  free(tmpv);
  /* dataprepFVT.f90:164-173: cell centre -> edge mid-point vectors */
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < 3; ++j) {
      int k = s->gedge[3 * i + j];
      double xm = HALF * (s->x[s->ep[2 * k]] + s->x[s->ep[2 * k + 1]]);
      double ym = HALF * (s->y[s->ep[2 * k]] + s->y[s->ep[2 * k + 1]]);
      s->rc[6 * i + 2 * j] = xm - s->xc[i];
      s->rc[6 * i + 2 * j + 1] = ym - s->yc[i];
    }
  /* dataprepFVT.f90:176-181: unit normals and tangents */
  for (int k = 0; k < E; ++k) {
    s->nrm[2 * k] = s->nrm[2 * k] / s->len[k];
    s->nrm[2 * k + 1] = s->nrm[2 * k + 1] / s->len[k];
    s->tng[2 * k] = s->tng[2 * k] / s->len[k];
    s->tng[2 * k + 1] = s->tng[2 * k + 1] / s->len[k];
  }
  /* lsq_setup.f90:92-145 (solver 2, weighted type 2), or with ACTIVATE HDLS
   * hdlsq_setup.f90:39-74 on t2tHD, unweighted (type 1, dataprepFVT.f90:219) */
  for (int i = 0; i < N; ++i) {
    double a = ZERO, b = ZERO, cc = ZERO;
    for (int t = s->t2t_ptr[i]; t < s->t2t_ptr[i + 1]; ++t) {
      int l = s->t2t[t];
      s->lsq_dx[t] = s->xc[l] - s->xc[i];
      s->lsq_dy[t] = s->yc[l] - s->yc[i];
      s->lsq_w[t] = s->o.gradient == ORC_GRAD_HDLS
                        ? ONE
                        : ONE / (s->lsq_dx[t] * s->lsq_dx[t] + s->lsq_dy[t] * s->lsq_dy[t]);
    }
    for (int t = s->t2t_ptr[i]; t < s->t2t_ptr[i + 1]; ++t) {
      a = a + s->lsq_w[t] * s->lsq_dx[t] * s->lsq_dx[t];
      b = b + s->lsq_w[t] * s->lsq_dx[t] * s->lsq_dy[t];
      cc = cc + s->lsq_w[t] * s->lsq_dy[t] * s->lsq_dy[t];
    }
    s->ai[i] = a; s->bi[i] = b; s->ci[i] = cc;
    s->det[i] = a * cc - b * b;
  }
  /* Green-Gauss gradient set-up on the triangle of neighbour centroids with
   * mirrored ghost points at boundary sides (gaussgrad_setup.f90:33-81,
   * ghost_cell.f90:34-41), in corrected form: neighbour centroids (not the
   * cell's own, :38), t2t3(j,i) (not t2t3(i,j), :35) and outward normals of
   * the auxiliary triangle, so that the gradient is exact for linear fields. */
  for (int i = 0; i < N; ++i) {
    double xp[3], yp[3];
    for (int j = 0; j < 3; ++j) {
      int k = s->t2t3[3 * i + j];
      if (k >= 0) {
        xp[j] = s->xc[k]; yp[j] = s->yc[k]; s->gconn[3 * i + j] = k;
      } else {
        int l = s->gedge[3 * i + j];
        double x1 = s->x[s->ep[2 * l]], y1 = s->y[s->ep[2 * l]];
        double x2 = s->x[s->ep[2 * l + 1]], y2 = s->y[s->ep[2 * l + 1]];
        double ax = s->xc[i] - x1, ay = s->yc[i] - y1, bx = x2 - x1, by = y2 - y1;
        double den = bx * bx + by * by;
        double cx = (bx * (ax * bx + ay * by) + by * (ay * bx - ax * by)) / den;
        double cy = (bx * (ax * by - ay * bx) + by * (ax * bx + ay * by)) / den;
        xp[j] = cx + x1; yp[j] = cy + y1; s->gconn[3 * i + j] = i;
      }
    }
    double ar = HALF * ((xp[1] - xp[0]) * (yp[2] - yp[0]) - (xp[2] - xp[0]) * (yp[1] - yp[0]));
    /* outward normals (for a counter-clockwise triangle) of sides 1-2, 2-3, 3-1 */
    double n1x = yp[1] - yp[0], n1y = xp[0] - xp[1];
    double n2x = yp[2] - yp[1], n2y = xp[1] - xp[2];
    double n3x = yp[0] - yp[2], n3y = xp[2] - xp[0];
    s->gcx[3 * i] = HALF * (n1x + n3x) / ar;
    s->gcx[3 * i + 1] = HALF * (n1x + n2x) / ar;
    s->gcx[3 * i + 2] = HALF * (n2x + n3x) / ar;
    s->gcy[3 * i] = HALF * (n1y + n3y) / ar;
    s->gcy[3 * i + 1] = HALF * (n1y + n2y) / ar;
    s->gcy[3 * i + 2] = HALF * (n2y + n3y) / ar;
    double den = (xp[1] - xp[0]) * (yp[2] - yp[1]) - (xp[2] - xp[1]) * (yp[1] - yp[0]);
    s->gwt[3 * i] = ((xp[1] - s->xc[i]) * (yp[2] - yp[1]) - (xp[2] - xp[1]) * (yp[1] - s->yc[i])) / den;
    s->gwt[3 * i + 1] = ((xp[2] - s->xc[i]) * (yp[0] - yp[2]) - (xp[0] - xp[2]) * (yp[2] - s->yc[i])) / den;
    s->gwt[3 * i + 2] = ((xp[0] - s->xc[i]) * (yp[1] - yp[0]) - (xp[1] - xp[0]) * (yp[0] - s->yc[i])) / den;
  }
  /* dataprepFVT.f90:256-273: SSP-RK Shu-Osher tables */
  memset(s->rka, 0, sizeof(s->rka));
  memset(s->rkb, 0, sizeof(s->rkb));
  switch (s->o.rkorder) {
    case 1: s->rka[0][0] = ONE; s->rkb[0] = ONE; break;
    case 2:
      s->rka[0][0] = ONE; s->rkb[0] = ONE;
      s->rka[1][0] = HALF; s->rka[1][1] = HALF; s->rkb[1] = HALF; break;
    case 3:
      s->rka[0][0] = ONE;
      s->rka[1][0] = 0.750; s->rka[1][1] = 0.250;
      s->rka[2][0] = ONE_THIRD; s->rka[2][1] = ZERO; s->rka[2][2] = 2.0 / 3.0;
      s->rkb[0] = ONE; s->rkb[1] = 0.250; s->rkb[2] = 2.0 / 3.0; break;
    default: return -1;
  }
  /* dataprepFVT.f90:278-286 */
  s->bch_max = ZERO;
  for (int l = 0; l < s->nbc; ++l) {
    if (!s->qbar[l]) continue;
    int k = s->bcedges[l];
    s->bch_max = fmax2(s->bch_max, s->zvtx[s->ep[2 * k]]);
    s->bch_max = fmax2(s->bch_max, s->zvtx[s->ep[2 * k + 1]]);
  }
  /* dataprepFVT.f90:293-306: vertices sorted by bed elevation */
  for (int i = 0; i < N; ++i) {
    int *p = &s->csortedz[3 * i];
    p[0] = s->vtx[3 * i]; p[1] = s->vtx[3 * i + 1]; p[2] = s->vtx[3 * i + 2];
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k)
        if (s->zvtx[p[k]] > s->zvtx[p[k + 1]]) { int a = p[k]; p[k] = p[k + 1]; p[k + 1] = a; }
  }
  return 0;
}

orc_t *orc_create(const orc_case_t *c, const orc_opts_t *o, char *err, int errlen) {
  char dummy[8];
  if (!err) { err = dummy; errlen = sizeof(dummy); }
  err[0] = 0;
  if (o->rkorder < 1 || o->rkorder > 3) { snprintf(err, errlen, "Incorrect Runge-Kutta order"); return NULL; }
  orc_t *s = xcalloc(1, sizeof(orc_t));
  s->P = c->n_pts; s->N = c->n_elems; s->E = c->n_edges;
  s->o = *o;
  const size_t P = s->P, N = s->N, E = s->E;
  s->x = xcalloc(P, 8); s->y = xcalloc(P, 8);
  s->vtx = xcalloc(3 * N, 4); s->gedge = xcalloc(3 * N, 4); s->gsign = xcalloc(3 * N, 4);
  s->area = xcalloc(N, 8); s->perim = xcalloc(N, 8); s->xc = xcalloc(N, 8); s->yc = xcalloc(N, 8);
  s->ep = xcalloc(2 * E, 4); s->ee = xcalloc(2 * E, 4);
  s->nrm = xcalloc(2 * E, 8); s->tng = xcalloc(2 * E, 8); s->len = xcalloc(E, 8);
  s->t2t3 = xcalloc(3 * N, 4);
  if (build(s, c, err, errlen) != 0) { orc_destroy(s); return NULL; }
  s->rc = xcalloc(6 * N, 8);
  s->c2v_w = xcalloc(s->n2t_ptr[P] + 1, 8); s->c2v_wid = xcalloc(s->n2t_ptr[P] + 1, 8);
  s->lsq_dx = xcalloc(s->t2t_ptr[N] + 1, 8); s->lsq_dy = xcalloc(s->t2t_ptr[N] + 1, 8);
  s->lsq_w = xcalloc(s->t2t_ptr[N] + 1, 8);
  s->ai = xcalloc(N, 8); s->bi = xcalloc(N, 8); s->ci = xcalloc(N, 8); s->det = xcalloc(N, 8);
  s->csortedz = xcalloc(3 * N, 4);
  s->gconn = xcalloc(3 * N, 4); s->gcx = xcalloc(3 * N, 8); s->gcy = xcalloc(3 * N, 8); s->gwt = xcalloc(3 * N, 8);
  s->z = xcalloc(N, 8); s->zvtx = xcalloc(P, 8); s->cd = xcalloc(N, 8);
  s->h = xcalloc(N + 1, 8); s->u = xcalloc(N, 8); s->v = xcalloc(N, 8);
  s->zeta = xcalloc(N, 8); s->u_mag = xcalloc(N, 8); s->wavec = xcalloc(N, 8); s->rough = xcalloc(N, 8);
  s->hvtx = xcalloc(P, 8); s->zetavtx = xcalloc(P, 8);
  s->delh = xcalloc(2 * N, 8); s->delu = xcalloc(2 * N, 8); s->delv = xcalloc(2 * N, 8);
  s->delz = xcalloc(2 * N, 8); s->gwork = xcalloc(2 * N, 8);
  s->phi = xcalloc(3 * N, 8);
  s->dry = xcalloc(N, 1);
  s->rk = xcalloc(3 * (size_t)(o->rkorder + 1) * N, 8);
  s->uvtx = xcalloc(P, 8); s->vvtx = xcalloc(P, 8); s->cdvtx = xcalloc(P, 8); s->e_est = xcalloc(N, 8);
  s->p_zeta = xcalloc(N, 8); s->p_cdvtx = xcalloc(P, 8); s->p_hvtx = xcalloc(P, 8);
  s->p_uvtx = xcalloc(P, 8); s->p_vvtx = xcalloc(P, 8); s->p_zetavtx = xcalloc(P, 8);
  if (dataprep(s, c) != 0) { snprintf(err, errlen, "dataprep failed"); orc_destroy(s); return NULL; }
  return s;
}

void orc_destroy(orc_t *s) {
  if (!s) return;
  void *ptrs[] = {s->x, s->y, s->vtx, s->gedge, s->gsign, s->area, s->perim, s->xc, s->yc, s->ep, s->ee,
                  s->nrm, s->tng, s->len, s->flowedg1, s->walledg1, s->bpolygon, s->bcedges, s->hbar, s->qbar,
                  s->in_type, s->in_ptr, s->in_edges, s->in_nodes_ptr, s->in_nodes, s->out_type, s->out_ptr,
                  s->out_edges, s->out_nodes_ptr, s->out_nodes, s->qts_ptr, s->hts_ptr, s->qts_t, s->qts_v,
                  s->hts_t, s->hts_v, s->qin, s->hbc, s->n2t_ptr, s->n2t, s->t2t_ptr, s->t2t, s->t2t3, s->rc,
                  s->c2v_w, s->c2v_wid, s->lsq_dx, s->lsq_dy, s->lsq_w, s->ai, s->bi, s->ci, s->det,
                  s->csortedz, s->gconn, s->gcx, s->gcy, s->gwt, s->z, s->zvtx, s->cd, s->h, s->u, s->v,
                  s->zeta, s->u_mag, s->wavec, s->rough, s->hvtx, s->zetavtx, s->delh, s->delu, s->delv,
                  s->delz, s->gwork, s->phi, s->dry, s->rk, s->pn, s->lin, s->hn, s->wall_pts,
                  s->wtang, s->uvtx, s->vvtx, s->cdvtx, s->e_est, s->p_zeta, s->p_cdvtx,
                  s->p_hvtx, s->p_uvtx, s->p_vvtx, s->p_zetavtx, s->et2t_ptr, s->et2t,
                  s->wcoef1, s->wcoef2, s->cvin, s->cvout, s->cvtab_ptr, s->cvpin, s->cvrin,
                  s->cvhead, s->cvq, s->cvsrc, s->phi0};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
  free(s);
}

/* ======================================================================
 * Numerical fluxes
 * ====================================================================== */

/* flux_ac.f90:48-126: Roe flux (Anastasiou & Chan 1997). */
static void flux_ac(double nx, double ny, double hlefty, double uleft, double vleft,
                    double hrighty, double uright, double vright, double *flux) {
  const double delta = 0.10;
  double hleft = hlefty, hright = hrighty;
  double fl[3], gl[3], fr[3], gr[3];
  fl[0] = hleft * uleft;
  fl[1] = hleft * uleft * uleft + HALF * G * hleft * hleft;
  fl[2] = hleft * uleft * vleft;
  gl[0] = hleft * vleft;
  gl[1] = hleft * uleft * vleft;
  gl[2] = hleft * vleft * vleft + HALF * G * hleft * hleft;
  fr[0] = hright * uright;
  fr[1] = hright * uright * uright + HALF * G * hright * hright;
  fr[2] = hright * uright * vright;
  gr[0] = hright * vright;
  gr[1] = hright * uright * vright;
  gr[2] = hright * vright * vright + HALF * G * hright * hright;
  hleft = fmax2(hleft, SMALL);
  double hls = sqrt(hleft);
  hright = fmax2(hright, SMALL);
  double hrs = sqrt(hright);
  double denom = hrs + hls;
  double uavg = (uright * hrs + uleft * hls) / denom;
  double vavg = (vright * hrs + vleft * hls) / denom;
  double cavg = sqrt(HALF * G * (hleft + hright));
  double q[3] = {hright - hleft, hright * uright - hleft * uleft, hright * vright - hleft * vleft};
  double rev[3][3], lev[3][3], eig[3], t1[3][3], t2[3][3], aq[3];
  rev[0][0] = ZERO; rev[0][1] = ONE; rev[0][2] = ONE;
  rev[1][0] = ny; rev[1][1] = uavg - cavg * nx; rev[1][2] = uavg + cavg * nx;
  rev[2][0] = -nx; rev[2][1] = vavg - cavg * ny; rev[2][2] = vavg + cavg * ny;
  lev[0][0] = -(uavg * ny - vavg * nx);
  lev[0][1] = ny;
  lev[0][2] = -nx;
  lev[1][0] = HALF * (uavg * nx + vavg * ny) / cavg + HALF;
  lev[1][1] = -(HALF * nx / cavg);
  lev[1][2] = -(HALF * ny / cavg);
  lev[2][0] = -(HALF * (uavg * nx + vavg * ny) / cavg) + HALF;
  lev[2][1] = HALF * nx / cavg;
  lev[2][2] = HALF * ny / cavg;
  eig[0] = fmax2(fabs(uavg * nx + vavg * ny), delta);
  eig[1] = fmax2(fabs(uavg * nx + vavg * ny - cavg), delta);
  eig[2] = fmax2(fabs(uavg * nx + vavg * ny + cavg), delta);
  /* MATMUL(rev, diag(eig)), MATMUL(., lev), MATMUL(., qdiff): sums in k order */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t1[i][j] = rev[i][j] * eig[j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t2[i][j] = t1[i][0] * lev[0][j] + t1[i][1] * lev[1][j] + t1[i][2] * lev[2][j];
  for (int i = 0; i < 3; ++i) aq[i] = t2[i][0] * q[0] + t2[i][1] * q[1] + t2[i][2] * q[2];
  for (int i = 0; i < 3; ++i)
    flux[i] = HALF * (fr[i] * nx + gr[i] * ny + fl[i] * nx + gl[i] * ny - aq[i]);
}

/* flux_rusanov.f90:47-80: local Lax-Friedrichs with caller-supplied
 * celerities. */
static void flux_rusanov(double nx, double ny, double hleft, double uleft, double vleft, double cl,
                         double hright, double uright, double vright, double cr, double *flux) {
  double fl[3], gl[3], fr[3], gr[3], q[3];
  fl[0] = hleft * uleft;
  fl[1] = hleft * uleft * uleft + HALF * G * hleft * hleft;
  fl[2] = hleft * uleft * vleft;
  gl[0] = hleft * vleft;
  gl[1] = hleft * uleft * vleft;
  gl[2] = hleft * vleft * vleft + HALF * G * hleft * hleft;
  fr[0] = hright * uright;
  fr[1] = hright * uright * uright + HALF * G * hright * hright;
  fr[2] = hright * uright * vright;
  gr[0] = hright * vright;
  gr[1] = hright * uright * vright;
  gr[2] = hright * vright * vright + HALF * G * hright * hright;
  double ul = uleft * nx + vleft * ny, ur = uright * nx + vright * ny;
  double splus = fmax2(fabs(ul) + cl, fabs(ur) + cr);
  q[0] = hright - hleft;
  q[1] = hright * uright - hleft * uleft;
  q[2] = hright * vright - hleft * vleft;
  for (int i = 0; i < 3; ++i)
    flux[i] = HALF * (fr[i] * nx + gr[i] * ny + fl[i] * nx + gl[i] * ny - splus * q[i]);
}

/* HLL (Toro 2001; Yoon & Kang 2004 eqs 6-8).  Wave-speed estimates with the
 * dry-bed variants of flux_HLL.f90:43-69; the star-region flux in corrected
 * form (flux_HLL.f90:103-109 omits the division by sR-sL on the flux terms,
 * overwrites flux(2) and never sets flux(3)). */
static void flux_hll(double nx, double ny, double hleft, double uleft, double vleft,
                     double hright, double uright, double vright, double h0, double *flux) {
  double cleft = sqrt(G * hleft), cright = sqrt(G * hright);
  double ustar = HALF * ((uleft + uright) * nx + (vleft + vright) * ny) + cleft - cright;
  double cstar = HALF * (cleft + cright) + ((uleft - uright) * nx + (vleft - vright) * ny) / 4.0;
  double sl, sr;
  if (hleft < h0) sl = uright * nx + vright * ny - 2.0 * cright;
  else if (hright < h0) sl = uleft * nx + vleft * ny - cleft;
  else sl = fmin2(uleft * nx + vleft * ny - cleft, ustar - cstar);
  if (hleft < h0) sr = uright * nx + vright * ny + cright;
  else if (hright < h0) sr = uleft * nx + vleft * ny + 2.0 * cleft;
  else sr = fmax2(uright * nx + vright * ny + cright, ustar + cstar);
  double FL[3], FR[3];
  FL[0] = hleft * uleft * nx + hleft * vleft * ny;
  FL[1] = (hleft * uleft * uleft + HALF * G * hleft * hleft) * nx + hleft * uleft * vleft * ny;
  FL[2] = hleft * uleft * vleft * nx + (hleft * vleft * vleft + HALF * G * hleft * hleft) * ny;
  FR[0] = hright * uright * nx + hright * vright * ny;
  FR[1] = (hright * uright * uright + HALF * G * hright * hright) * nx + hright * uright * vright * ny;
  FR[2] = hright * uright * vright * nx + (hright * vright * vright + HALF * G * hright * hright) * ny;
  if (sl >= ZERO) {
    flux[0] = FL[0]; flux[1] = FL[1]; flux[2] = FL[2];
  } else if (sr <= ZERO) {
    flux[0] = FR[0]; flux[1] = FR[1]; flux[2] = FR[2];
  } else {
    double d = sr - sl;
    double q[3] = {hright - hleft, hright * uright - hleft * uleft, hright * vright - hleft * vleft};
    for (int i = 0; i < 3; ++i) flux[i] = (sr * FL[i] - sl * FR[i] + sl * sr * q[i]) / d;
  }
}

/* Roe flux of Brufau & Garcia-Navarro with pre-projected physical fluxes
 * (flux_bgnvc.f90:41-103), with the left eigenvectors corrected to the
 * inverse of the right ones: L(1,1) = c - u.n (flux_bgnvc.f90:79 has
 * u.n + c) and L(3,3) = -ny (:87 has vavg - ny). */
static void flux_bgnvc(double nx, double ny, double hleft, double uleft, double vleft,
                       double hright, double uright, double vright, double *flux) {
  double fl[3], gl[3], fr[3], gr[3], q[3];
  fl[0] = hleft * uleft * nx;
  fl[1] = (hleft * uleft * uleft + HALF * G * hleft * hleft) * nx;
  fl[2] = hleft * uleft * vleft * nx;
  gl[0] = hleft * vleft * ny;
  gl[1] = hleft * uleft * vleft * ny;
  gl[2] = (hleft * vleft * vleft + HALF * G * hleft * hleft) * ny;
  fr[0] = hright * uright * nx;
  fr[1] = (hright * uright * uright + HALF * G * hright * hright) * nx;
  fr[2] = hright * uright * vright * nx;
  gr[0] = hright * vright * ny;
  gr[1] = hright * uright * vright * ny;
  gr[2] = (hright * vright * vright + HALF * G * hright * hright) * ny;
  double hls = sqrt(hleft), hrs = sqrt(hright);
  double denom = hrs + hls;
  double uavg = (uright * hrs + uleft * hls) / denom;
  double vavg = (vright * hrs + vleft * hls) / denom;
  double cavg = sqrt(HALF * G * (hleft + hright));
  q[0] = hright - hleft;
  q[1] = hright * uright - hleft * uleft;
  q[2] = hright * vright - hleft * vleft;
  double un = uavg * nx + vavg * ny;
  double rev[3][3] = {{ONE, ZERO, ONE},
                      {uavg + cavg * nx, -cavg * ny, uavg - cavg * nx},
                      {vavg + cavg * ny, cavg * nx, vavg - cavg * ny}};
  double lev[3][3] = {{cavg - un, nx, ny},
                      {2.0 * (uavg * ny - vavg * nx), -2.0 * ny, 2.0 * nx},
                      {un + cavg, -nx, -ny}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) lev[i][j] = HALF * lev[i][j] / cavg;
  double eig[3] = {fabs(un + cavg), fabs(un), fabs(un - cavg)};
  double t1[3][3], t2[3][3], aq[3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t1[i][j] = rev[i][j] * eig[j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t2[i][j] = t1[i][0] * lev[0][j] + t1[i][1] * lev[1][j] + t1[i][2] * lev[2][j];
  for (int i = 0; i < 3; ++i) aq[i] = t2[i][0] * q[0] + t2[i][1] * q[1] + t2[i][2] * q[2];
  for (int i = 0; i < 3; ++i) flux[i] = HALF * (fr[i] + gr[i] + fl[i] + gl[i] - aq[i]);
}

void orc_flux(int32_t kind, double nx, double ny, double hl, double ul, double vl,
              double hr, double ur, double vr, double cl, double cr, double h0, double *f) {
  switch (kind) {
    case ORC_FLUX_AC: flux_ac(nx, ny, hl, ul, vl, hr, ur, vr, f); break;
    case ORC_FLUX_HLL: flux_hll(nx, ny, hl, ul, vl, hr, ur, vr, h0, f); break;
    case ORC_FLUX_RUSANOV: flux_rusanov(nx, ny, hl, ul, vl, cl, hr, ur, vr, cr, f); break;
    default: flux_bgnvc(nx, ny, hl, ul, vl, hr, ur, vr, f); break;
  }
}

/* ======================================================================
 * Stage routines
 * ====================================================================== */

/* find_dry_cells.f90:25-42 (+ wavec, solverFVT.f90:180-181) */
static void find_dry_cells(orc_t *s, int with_wavec) {
  for (int i = 0; i < s->N; ++i) s->dry[i] = s->h[i] < s->o.h_dry;
  for (int i = 0; i < s->N; ++i) {
    s->u_mag[i] = ZERO;
    if (!s->dry[i]) s->u_mag[i] = sqrt(s->u[i] * s->u[i] + s->v[i] * s->v[i]);
  }
  for (int i = 0; i < s->N; ++i) {
    if (s->dry[i]) continue;
    if (s->h[i] < s->o.h_wet && orc_equals(s->u_mag[i], ZERO)) s->dry[i] = 1;
  }
  if (with_wavec)
    for (int i = 0; i < s->N; ++i) s->wavec[i] = s->dry[i] ? ZERO : sqrt(G * s->h[i]);
}

/* lsq_grad_array.f90:60-73; on t2tHD with unit weights the same sums are
 * hdlsq_grad_array.f90:36-48 */
static void lsq_grad(const orc_t *s, const double *ph, double *gr) {
  for (int i = 0; i < s->N; ++i) {
    double di = ZERO, ei = ZERO;
    for (int t = s->t2t_ptr[i]; t < s->t2t_ptr[i + 1]; ++t) {
      double dphi = ph[s->t2t[t]] - ph[i];
      di = di + s->lsq_w[t] * dphi * s->lsq_dx[t];
      ei = ei + s->lsq_w[t] * dphi * s->lsq_dy[t];
    }
    gr[2 * i] = (s->ci[i] * di - s->bi[i] * ei) / s->det[i];
    gr[2 * i + 1] = (s->ai[i] * ei - di * s->bi[i]) / s->det[i];
  }
}

/* gaussgrad.f90:37-52 */
static void gauss_grad(orc_t *s, const double *ph, double *gr) {
  for (int i = 0; i < s->N; ++i) {
    const int *g = &s->gconn[3 * i];
    s->gwork[2 * i] = s->gcx[3 * i] * ph[g[0]] + s->gcx[3 * i + 1] * ph[g[1]] + s->gcx[3 * i + 2] * ph[g[2]];
    s->gwork[2 * i + 1] = s->gcy[3 * i] * ph[g[0]] + s->gcy[3 * i + 1] * ph[g[1]] + s->gcy[3 * i + 2] * ph[g[2]];
  }
  for (int i = 0; i < s->N; ++i) {
    const int *g = &s->gconn[3 * i];
    const double *w = &s->gwt[3 * i];
    gr[2 * i] = HALF * (s->gwork[2 * i] + w[0] * s->gwork[2 * g[0]] + w[1] * s->gwork[2 * g[1]] + w[2] * s->gwork[2 * g[2]]);
    gr[2 * i + 1] = HALF * (s->gwork[2 * i + 1] + w[0] * s->gwork[2 * g[0] + 1] + w[1] * s->gwork[2 * g[1] + 1] + w[2] * s->gwork[2 * g[2] + 1]);
  }
}

/* limiterBJ.f90:42-81 (beta = 1, cell_grads.f90:59) */
static void limiter_bj(const orc_t *s, double beta, const double *q, double *gq) {
  for (int i = 0; i < s->N; ++i) {
    double q0 = q[i], r[3];
    for (int j = 0; j < 3; ++j) {
      int k = s->t2t3[3 * i + j];
      double qn = (k < 0) ? q0 : q[k];
      double qmin = fmin2(q0, qn), qmax = fmax2(q0, qn);
      double qj = q0 + gq[2 * i] * s->rc[6 * i + 2 * j] + gq[2 * i + 1] * s->rc[6 * i + 2 * j + 1];
      double dq = qj - q0;
      if (orc_equals(dq, ZERO)) r[j] = ONE;
      else if (dq < 0) r[j] = (qmin - q0) / dq;
      else r[j] = (qmax - q0) / dq;
      double a = fmin2(beta * r[j], ONE);
      double b = fmin2(r[j], beta);
      r[j] = fmax2(a, b);
    }
    double ph = fmin2(fmin2(r[0], r[1]), r[2]);
    gq[2 * i] = ph * gq[2 * i];
    gq[2 * i + 1] = ph * gq[2 * i + 1];
  }
}

/* cell_grads.f90:24-75 */
static void cell_grads(orc_t *s) {
  const int N = s->N;
  if (s->o.limiter == 0) {
    memset(s->delh, 0, sizeof(double) * 2 * N);
    memset(s->delu, 0, sizeof(double) * 2 * N);
    memset(s->delv, 0, sizeof(double) * 2 * N);
  } else {
    if (s->o.gradient == ORC_GRAD_GAUSS) {
      gauss_grad(s, s->zeta, s->delh);
      gauss_grad(s, s->u, s->delu);
      gauss_grad(s, s->v, s->delv);
    } else {
      lsq_grad(s, s->zeta, s->delh);
      lsq_grad(s, s->u, s->delu);
      lsq_grad(s, s->v, s->delv);
    }
    if (s->o.limiter == 2) {
      limiter_bj(s, ONE, s->zeta, s->delh);
      limiter_bj(s, ONE, s->u, s->delu);
      limiter_bj(s, ONE, s->v, s->delv);
    }
  }
  for (int i = 0; i < N; ++i) {
    s->delh[2 * i] = s->delh[2 * i] - s->delz[2 * i];
    s->delh[2 * i + 1] = s->delh[2 * i + 1] - s->delz[2 * i + 1];
  }
}

/* hgradfix2.f90:46-98 (solver 2) */
static void hgradfix2(orc_t *s, int rkstep) {
  for (int i = 0; i < s->N; ++i) {
    if (s->dry[i]) continue;
    int setgrad = s->zeta[i] < s->zvtx[s->csortedz[3 * i + 2]];
    for (int k = 0; k < 3; ++k) {
      int l = s->t2t3[3 * i + k];
      if (l < 0) continue;
      if (s->dry[l] && s->z[l] < s->zeta[i]) setgrad = 0;
    }
    if (setgrad) {
      s->delh[2 * i] = -s->delz[2 * i];
      s->delh[2 * i + 1] = -s->delz[2 * i + 1];
      s->u[i] = ZERO;
      s->v[i] = ZERO;
      RK(s, 1, rkstep - 1, i) = ZERO;
      RK(s, 2, rkstep - 1, i) = ZERO;
      s->delu[2 * i] = ZERO; s->delu[2 * i + 1] = ZERO;
      s->delv[2 * i] = ZERO; s->delv[2 * i + 1] = ZERO;
    }
  }
}

/* enforce_bcs.f90:37-56, the only part that feeds the cell state: the stage
 * prescribed at the nodes of stage (htype 1) strings. */
static void enforce_bcs_h(orc_t *s) {
  for (int ch = 0; ch < s->n_out; ++ch) {
    if (s->out_type[ch] != 1) continue;
    for (int t = s->out_nodes_ptr[ch]; t < s->out_nodes_ptr[ch + 1]; ++t) {
      int j = s->out_nodes[t];
      s->hvtx[j] = fmax2(ZERO, s->hbc[ch] - s->zvtx[j]);
    }
  }
}

/* to_vertices.f90:26-55 with ctr2vtxWSE.f90:89-105 (type 3) and
 * positivity_vtx.f90:22-27: vertex water surface and depth, and the cell
 * velocities zeroed in partially wet cells. */
static void to_vertices_wse(orc_t *s) {
  for (int i = 0; i < s->P; ++i) {
    double id = ZERO, a = ZERO;
    int counter = 0;
    for (int j = s->n2t_ptr[i]; j < s->n2t_ptr[i + 1]; ++j) {
      int k = s->n2t[j];
      if (s->dry[k]) continue;
      counter++;
      id = id + s->c2v_wid[j];
      a = a + s->c2v_wid[j] * s->zeta[k];
    }
    s->zetavtx[i] = (counter > 0) ? a / id : s->zvtx[i];
  }
  for (int i = 0; i < s->P; ++i) s->hvtx[i] = s->zetavtx[i] - s->zvtx[i];
  for (int i = 0; i < s->P; ++i) {
    double zr = s->zvtx[i];
    for (int j = s->n2t_ptr[i]; j < s->n2t_ptr[i + 1]; ++j) {
      int k = s->n2t[j];
      if (s->dry[k]) continue;
      zr = fmax2(s->zeta[k], zr);
    }
    s->zetavtx[i] = fmin2(s->zetavtx[i], zr);
    s->hvtx[i] = s->zetavtx[i] - s->zvtx[i];
  }
  for (int i = 0; i < s->P; ++i)
    if (s->hvtx[i] < ZERO) { s->hvtx[i] = ZERO; s->zetavtx[i] = s->zvtx[i]; }
  for (int i = 0; i < s->N; ++i) {
    if (s->dry[i]) continue;
    int flag = 1;
    for (int j = 0; j < 3; ++j) flag = flag && (s->hvtx[s->vtx[3 * i + j]] > s->o.h_dry);
    if (!flag) { s->u[i] = ZERO; s->v[i] = ZERO; }
  }
}

/* to_vertices as every RK stage calls it: the vertex velocities it also
 * forms (:56-64) are output only, so of enforce_bcs only the stage strings'
 * depths (:37-56), which the wall fluxes read, are formed here. */
static void to_vertices(orc_t *s) {
  to_vertices_wse(s);
  enforce_bcs_h(s);
}

/* vbc_by_h.f90:32-80 for inflow string ch on vertex depths h0: the node
 * speeds v_inflow (zero at the string's end nodes). */
static void vbc_by_h(const orc_t *s, int ch, const double *h0, double *vin) {
  const int a = s->in_nodes_ptr[ch], n = s->in_nodes_ptr[ch + 1] - a;
  const double qin = s->qin[ch], h_dry = s->o.h_dry;
  double *hi = xcalloc(n, 8), *ai = xcalloc(n, 8), *ui = xcalloc(n, 8);
  for (int i = 0; i < n; ++i) hi[i] = h0[s->in_nodes[a + i]];
  double ht = ZERO, at = ZERO;
  for (int i = 0; i + 1 < n; ++i) {
    ai[i] = s->lin[a + i] * (hi[i] + hi[i + 1]) * HALF;
    ht = ht + (hi[i] + hi[i + 1]) * HALF;
    at = at + ai[i];
  }
  for (int i = 0; i + 1 < n; ++i) ui[i] = HALF * (hi[i] + hi[i + 1]) * qin / ht / at;
  for (int i = 0; i < n; ++i) vin[i] = ZERO;
  for (int i = 1; i + 1 < n; ++i)
    if (hi[i] > h_dry) vin[i] = HALF * (ui[i - 1] + ui[i]);
  for (int i = 0; i + 1 < n; ++i) ui[i] = HALF * (vin[i] + vin[i + 1]);
  double q = ZERO;
  for (int i = 0; i + 1 < n; ++i) q = q + ui[i] * ai[i];
  for (int i = 1; i + 1 < n; ++i)
    if (hi[i] > h_dry) vin[i] = vin[i] * qin / q;
  free(hi); free(ai); free(ui);
}

/* enforce_bcs.f90:30-171 on the vertex fields (hvtx, uvtx, vvtx, zvtx). */
static void enforce_bcs(orc_t *s) {
  double *h0 = s->hvtx, *u0 = s->uvtx, *v0 = s->vvtx;
  const double h_dry = s->o.h_dry;
  for (int ch = 0; ch < s->n_out; ++ch) {
    const int a = s->out_nodes_ptr[ch], b = s->out_nodes_ptr[ch + 1];
    const int ht = s->out_type[ch];
    if (ht == 1)
      for (int t = a; t < b; ++t) {
        const int j = s->out_nodes[t];
        h0[j] = fmax2(ZERO, s->hbc[ch] - s->zvtx[j]);
      }
    if (ht == 0 || (ht == 1 && s->o.hvalve))
      for (int t = a; t < b; ++t) {
        const int j = s->out_nodes[t];
        const double dot = u0[j] * s->hn[2 * t] + v0[j] * s->hn[2 * t + 1];
        if (dot < ZERO) { u0[j] = ZERO; v0[j] = ZERO; }
      }
  }
  if (s->o.btype == 0) {
    for (int i = 0; i < s->n_wall; ++i) { u0[s->wall_pts[i]] = ZERO; v0[s->wall_pts[i]] = ZERO; }
  } else {
    for (int i = 0; i < s->n_wall; ++i) {
      const int p = s->wall_pts[i];
      const double tx = s->wtang[2 * i], ty = s->wtang[2 * i + 1];
      const double dot = u0[p] * tx + v0[p] * ty;
      u0[p] = dot * tx;
      v0[p] = dot * ty;
    }
  }
  for (int i = 0; i < s->P; ++i)
    if (h0[i] < h_dry) { u0[i] = ZERO; v0[i] = ZERO; }
  for (int ch = 0; ch < s->n_in; ++ch) {
    const int a = s->in_nodes_ptr[ch], n = s->in_nodes_ptr[ch + 1] - a;
    const int vt = s->in_type[ch];
    if (vt != 0 && vt != 1 && vt != 5) continue;
    double *vin = xcalloc(n, 8);
    if (vt == 0) for (int i = 0; i < n; ++i) vin[i] = s->qin[ch];
    else vbc_by_h(s, ch, h0, vin);
    for (int i = 0; i < n; ++i) {
      const int j = s->in_nodes[a + i];
      if (h0[j] > h_dry) { u0[j] = vin[i] * s->pn[2 * (a + i)]; v0[j] = vin[i] * s->pn[2 * (a + i) + 1]; }
    }
    free(vin);
  }
}

/* The to_vertices call that ends solverFVT (solverFVT.f90:483): the water
 * surface and depth of to_vertices.f90:26-55, then the vertex velocities of
 * :56-65 (ctr2vtx of u and v, wetdryvfix.f90:30-49 at no-slip walls, zero at
 * dry vertices) and the whole of enforce_bcs. */
static void to_vertices_final(orc_t *s) {
  to_vertices_wse(s);
  ctr2vtx3(s, s->u, s->uvtx);
  ctr2vtx3(s, s->v, s->vvtx);
  if (s->o.btype == 0)
    for (int i = 0; i < s->P; ++i) {
      int wett = 0, dryt = 0;
      for (int j = s->n2t_ptr[i]; j < s->n2t_ptr[i + 1]; ++j) {
        if (s->dry[s->n2t[j]]) dryt = 1; else wett = 1;
      }
      if (wett && dryt) { s->uvtx[i] = ZERO; s->vvtx[i] = ZERO; }
    }
  for (int i = 0; i < s->P; ++i)
    if (s->hvtx[i] < s->o.h_dry) { s->uvtx[i] = ZERO; s->vvtx[i] = ZERO; }
  enforce_bcs(s);
}

/* error_est.f90:57-84 (etype 2) */
static void error_est2(orc_t *s) {
  for (int i = 0; i < s->N; ++i) {
    s->e_est[i] = ZERO;
    if (s->h[i] < s->o.h_dry) continue;
    int n1 = s->vtx[3 * i], n2 = s->vtx[3 * i + 1], n3 = s->vtx[3 * i + 2];
    double gzx = (s->zvtx[n1] * (s->y[n2] - s->y[n3]) + s->zvtx[n2] * (s->y[n3] - s->y[n1]) +
                  s->zvtx[n3] * (s->y[n1] - s->y[n2])) * HALF / s->area[i];
    double gzy = (s->zvtx[n1] * (s->x[n3] - s->x[n2]) + s->zvtx[n2] * (s->x[n1] - s->x[n3]) +
                  s->zvtx[n3] * (s->x[n2] - s->x[n1])) * HALF / s->area[i];
    double l = s->perim[i] * ONE_THIRD;
    double gx = gzx + s->delh[2 * i], gy = gzy + s->delh[2 * i + 1];
    s->e_est[i] = sqrt(gx * gx + gy * gy) * l / (s->z[i] + s->h[i]);
  }
}

static inline void face(const orc_t *s, int i, int m, double *h, double *u, double *v) {
  const double rx = s->rc[6 * i + 2 * m], ry = s->rc[6 * i + 2 * m + 1];
  *h = s->h[i] + s->delh[2 * i] * rx + s->delh[2 * i + 1] * ry;
  *u = s->u[i] + s->delu[2 * i] * rx + s->delu[2 * i + 1] * ry;
  *v = s->v[i] + s->delv[2 * i] * rx + s->delv[2 * i + 1] * ry;
}

/* invisc_fluxes.f90:39-186: interior edges.  With fluxes = 1, the loop of
 * invisc_fluxes_xsf.f90:41-164 instead: a dry side keeps its cell depth
 * (:58,74); the left state takes the y-components of the RIGHT cell's u and
 * v gradients (delu(j)%y, delv(j)%y at :65-66, kept as written); where
 * ALGORITHM I1 sets a wall, flux_ac runs on the unchanged states and only
 * the mass flux is zeroed (:121,125); the flux is always flux_ac. */
static void interior_fluxes(orc_t *s) {
  const double h_dry = s->o.h_dry, h_wet = s->o.h_wet;
  const int xsf = s->o.fluxes == 1;
  for (int l = 0; l < s->nflow1; ++l) {
    int k = s->flowedg1[l];
    int i = s->ee[2 * k], j = s->ee[2 * k + 1];
    int set_wall = 0;
    if (s->dry[i] && s->dry[j]) continue;
    int m = local_edge(s, i, k);
    double nx = s->nrm[2 * k], ny = s->nrm[2 * k + 1];
    if (s->gsign[3 * i + m] < 0) { nx = -nx; ny = -ny; }
    double hl, ul, vl, hr, ur, vr;
    if (s->dry[i]) { ul = ZERO; vl = ZERO; hl = xsf ? s->h[i] : ZERO; }
    else {
      face(s, i, m, &hl, &ul, &vl);
      if (hl < ZERO) hl = ZERO;
      if (xsf) {
        const double rx = s->rc[6 * i + 2 * m], ry = s->rc[6 * i + 2 * m + 1];
        ul = s->u[i] + s->delu[2 * i] * rx + s->delu[2 * j + 1] * ry;
        vl = s->v[i] + s->delv[2 * i] * rx + s->delv[2 * j + 1] * ry;
      }
    }
    int n = local_edge(s, j, k);
    if (s->dry[j]) { ur = ZERO; vr = ZERO; hr = xsf ? s->h[j] : ZERO; }
    else {
      face(s, j, n, &hr, &ur, &vr);
      if (hr < ZERO) hr = ZERO;
    }
    if (hl < h_dry && hr < h_dry) continue;
    if (s->dry[i] || s->dry[j]) {
      double h0 = fmax2(hr, hl);
      double z0 = HALF * (s->zvtx[s->ep[2 * k]] + s->zvtx[s->ep[2 * k + 1]]);
      h0 = h0 + z0;
      if (s->dry[i] && s->z[i] + h_wet > h0) { set_wall = 1; if (!xsf) hl = hr; }
      if (s->dry[j] && s->z[j] + h_wet > h0) { set_wall = 1; if (!xsf) hr = hl; }
    }
    double f[3];
    if (xsf) {
      flux_ac(nx, ny, hl, ul, vl, hr, ur, vr, f);
      if (set_wall) f[0] = ZERO;
    } else if (set_wall) {
      f[0] = ZERO;
      hl = fmax2(hl, hr);
      f[1] = HALF * G * hl * hl * nx;
      f[2] = HALF * G * hl * hl * ny;
    } else {
      switch (s->o.flux) {
        case ORC_FLUX_HLL: flux_hll(nx, ny, hl, ul, vl, hr, ur, vr, h_dry, f); break;
        case ORC_FLUX_RUSANOV: flux_rusanov(nx, ny, hl, ul, vl, s->wavec[i], hr, ur, vr, s->wavec[j], f); break;
        case ORC_FLUX_BGNVC: flux_bgnvc(nx, ny, hl, ul, vl, hr, ur, vr, f); break;
        default: flux_ac(nx, ny, hl, ul, vl, hr, ur, vr, f); break;
      }
    }
    const double L = s->len[k];
    s->phi[i] = s->phi[i] - f[0] * L;
    s->phi[s->N + i] = s->phi[s->N + i] - f[1] * L;
    s->phi[2 * s->N + i] = s->phi[2 * s->N + i] - f[2] * L;
    s->phi[j] = s->phi[j] + f[0] * L;
    s->phi[s->N + j] = s->phi[s->N + j] + f[1] * L;
    s->phi[2 * s->N + j] = s->phi[2 * s->N + j] + f[2] * L;
  }
}

/* wenslp3_fluxes.f90:34-67 (no-slip walls, btype 0) */
static void wenslp3(orc_t *s) {
  for (int l = 0; l < s->nwall1; ++l) {
    int k = s->walledg1[l];
    int i = edge_in_element(s, k);
    if (s->dry[i]) continue;
    double h1 = s->hvtx[s->ep[2 * k]], h2 = s->hvtx[s->ep[2 * k + 1]];
    double he = sqrt((h1 * h1 + h1 * h2 + h2 * h2) * ONE_THIRD);
    if (he < ZERO) he = ZERO;
    if (he < s->o.h_dry) continue;
    int m = local_edge(s, i, k);
    double nx = s->nrm[2 * k], ny = s->nrm[2 * k + 1];
    if (s->gsign[3 * i + m] < 0) { nx = -nx; ny = -ny; }
    s->phi[s->N + i] = s->phi[s->N + i] - HALF * G * he * he * nx * s->len[k];
    s->phi[2 * s->N + i] = s->phi[2 * s->N + i] - HALF * G * he * he * ny * s->len[k];
  }
}

/* weslp2_fluxes.f90:34-76 (slip walls, btype 1) */
static void weslp2(orc_t *s) {
  for (int l = 0; l < s->nwall1; ++l) {
    int k = s->walledg1[l];
    int i = edge_in_element(s, k);
    if (s->dry[i]) continue;
    int m = local_edge(s, i, k);
    double hl, ul, vl;
    face(s, i, m, &hl, &ul, &vl);
    if (hl < s->o.h_dry) continue;
    double nx = s->nrm[2 * k], ny = s->nrm[2 * k + 1], tx = s->tng[2 * k], ty = s->tng[2 * k + 1];
    if (s->gsign[3 * i + m] < 0) { nx = -nx; ny = -ny; tx = -tx; ty = -ty; }
    double ur = ul * tx + vl * ty;
    ur = ur + 2.0 * sqrt(G * hl);
    double hr = ur * ur / 4.0 / G;
    double f1 = HALF * G * hr * hr * nx, f2 = HALF * G * hr * hr * ny;
    double dh = ZERO;
    s->phi[s->N + i] = s->phi[s->N + i] - (f1 + dh * nx) * s->len[k];
    s->phi[2 * s->N + i] = s->phi[2 * s->N + i] - (f2 + dh * ny) * s->len[k];
  }
}

/* inflowQbyH.f90:31-212 (vtype 1) and :221-402 (vtype 5, which skips dry
 * cells).  One Riemann-invariant fixed point per string. */
static void inflow_q(orc_t *s, int vt) {
  const double h_dry = s->o.h_dry;
  for (int ch = 0; ch < s->n_in; ++ch) {
    if (s->in_type[ch] != vt) continue;
    if (s->qin[ch] < DBL_MIN) continue;
    const int e0 = s->in_ptr[ch], e1 = s->in_ptr[ch + 1];
    double atotal = ZERO, newatotal = ZERO, qtotal;
    for (int l = e0; l < e1; ++l) {
      int k = s->in_edges[l];
      int i = edge_in_element(s, k);
      int m = local_edge(s, i, k);
      if (vt == 5 && s->dry[i]) continue;
      double hl;
      if (s->dry[i]) hl = ZERO;
      else {
        hl = s->h[i] + s->delh[2 * i] * s->rc[6 * i + 2 * m] + s->delh[2 * i + 1] * s->rc[6 * i + 2 * m + 1];
        hl = fmax2(hl, ZERO);
      }
      atotal = atotal + hl * s->len[k];
    }
    int null_area = orc_equals(atotal, ZERO);
    int counter = 0;
    for (int pass = 0;; ++pass) {
      int final = (pass < 0);
      (void)final;
      newatotal = ZERO;
      qtotal = ZERO;
      for (int l = e0; l < e1; ++l) {
        int k = s->in_edges[l];
        int i = edge_in_element(s, k);
        if (vt == 5 && s->dry[i]) continue;
        int m = local_edge(s, i, k);
        double hl, ue, ve, tmpu, tmpv;
        face(s, i, m, &hl, &tmpu, &tmpv);
        if (null_area) hl = s->bch_max;
        if (hl < h_dry) { hl = fmax2(hl, ZERO); ue = ZERO; ve = ZERO; }
        else { ue = tmpu; ve = tmpv; }
        double nx = s->nrm[2 * k], ny = s->nrm[2 * k + 1];
        if (s->gsign[3 * i + m] < 0) { nx = -nx; ny = -ny; }
        double qfrac = s->qin[ch] / atotal;
        double ur = -qfrac;
        double ul = ue * nx + ve * ny;
        double sq = sqrt(G * hl);
        double hr = 2.0 * sq + ul - ur;
        hr = hr * hr / 4.0 / G;
        ur = -qfrac * nx;
        double vr = -qfrac * ny;
        if (hr < h_dry) continue;
        newatotal = newatotal + hr * s->len[k];
        qtotal = qtotal + (ur * nx + vr * ny) * hr * s->len[k];
      }
      qtotal = fabs(qtotal);
      if (fabs(qtotal - s->qin[ch]) / s->qin[ch] < 0.005 || counter > 50) break;
      counter = counter + 1;
      atotal = newatotal;
    }
    atotal = newatotal;
    for (int l = e0; l < e1; ++l) {
      int k = s->in_edges[l];
      int i = edge_in_element(s, k);
      if (vt == 5 && s->dry[i]) continue;
      int m = local_edge(s, i, k);
      double hl, ue, ve, tmpu, tmpv;
      face(s, i, m, &hl, &tmpu, &tmpv);
      if (null_area) hl = s->bch_max;
      if (hl < h_dry) { hl = fmax2(hl, ZERO); ue = ZERO; ve = ZERO; }
      else { ue = tmpu; ve = tmpv; }
      double nx = s->nrm[2 * k], ny = s->nrm[2 * k + 1];
      if (s->gsign[3 * i + m] < 0) { nx = -nx; ny = -ny; }
      double qfrac = s->qin[ch] / atotal;
      double ur = -qfrac;
      double ul = ue * nx + ve * ny;
      double sq = sqrt(G * hl);
      double hr = 2.0 * sq + ul - ur;
      hr = hr * hr / 4.0 / G;
      ur = -qfrac * nx;
      double vr = -qfrac * ny;
      if (hr < h_dry) continue;
      double f[3];
      f[0] = hr * (ur * nx + vr * ny);
      f[1] = hr * ur * (ur * nx + vr * ny) + HALF * G * hr * hr * nx;
      f[2] = hr * vr * (ur * nx + vr * ny) + HALF * G * hr * hr * ny;
      const double L = s->len[k];
      s->phi[i] = s->phi[i] - f[0] * L;
      s->phi[s->N + i] = s->phi[s->N + i] - f[1] * L;
      s->phi[2 * s->N + i] = s->phi[2 * s->N + i] - f[2] * L;
    }
  }
}

/* inflowH.f90:28-93 (vtype 4: stage prescribed at an inflow string) */
static void inflow_h(orc_t *s) {
  const double h_dry = s->o.h_dry;
  for (int ch = 0; ch < s->n_in; ++ch) {
    if (s->in_type[ch] != 4) continue;
    for (int l = s->in_ptr[ch]; l < s->in_ptr[ch + 1]; ++l) {
      int k = s->in_edges[l];
      int i = edge_in_element(s, k);
      int m = local_edge(s, i, k);
      double hl, ue, ve, tmpu, tmpv;
      face(s, i, m, &hl, &tmpu, &tmpv);
      hl = fmax2(hl, ZERO);
      if (s->h[i] < h_dry) { ue = ZERO; ve = ZERO; } else { ue = tmpu; ve = tmpv; }
      double nx = s->nrm[2 * k], ny = s->nrm[2 * k + 1];
      if (s->gsign[3 * i + m] < 0) { nx = -nx; ny = -ny; }
      double zm = HALF * (s->zvtx[s->ep[2 * k]] + s->zvtx[s->ep[2 * k + 1]]);
      double hr = fmax2(ZERO, s->qin[ch] - zm);
      double ul = ue * nx + ve * ny;
      double sl = sqrt(G * hl), sr = sqrt(G * hr);
      double ur = ul + (sl - sr);
      double vr = ur * ny;
      ur = ur * nx;
      if (hl < h_dry && hr < h_dry) continue;
      double f[3];
      flux_ac(nx, ny, hl, ue, ve, hr, ur, vr, f);
      for (int n = 0; n < 3; ++n) s->phi[n * s->N + i] = s->phi[n * s->N + i] - f[n] * s->len[k];
    }
  }
}

/* outflowbyH.f90:30-137 (htype 1: stage prescribed) */
static void outflow_h(orc_t *s) {
  const double h_dry = s->o.h_dry;
  for (int ch = 0; ch < s->n_out; ++ch) {
    if (s->out_type[ch] != 1) continue;
    for (int l = s->out_ptr[ch]; l < s->out_ptr[ch + 1]; ++l) {
      int k = s->out_edges[l];
      int i = edge_in_element(s, k);
      int m = local_edge(s, i, k);
      double hl, ue, ve, tmpu, tmpv;
      face(s, i, m, &hl, &tmpu, &tmpv);
      if (s->dry[i] || hl < h_dry) { ue = ZERO; ve = ZERO; hl = fmax2(hl, ZERO); }
      else { ue = tmpu; ve = tmpv; }
      double nx = s->nrm[2 * k], ny = s->nrm[2 * k + 1], tx = s->tng[2 * k], ty = s->tng[2 * k + 1];
      if (s->gsign[3 * i + m] < 0) { nx = -nx; ny = -ny; tx = -tx; ty = -ty; }
      double zm = HALF * (s->zvtx[s->ep[2 * k]] + s->zvtx[s->ep[2 * k + 1]]);
      double hr = fmax2(ZERO, s->hbc[ch] - zm);
      if (hl < h_dry && hr < h_dry) continue;
      double ul, vl, ur, vr;
      if (hr < h_dry) { ur = ZERO; ul = ue; vr = ZERO; vl = ve; }
      else {
        ul = ue * nx + ve * ny;
        vl = ue * tx + ve * ty;
        double sl = sqrt(G * hl), sr = sqrt(G * hr);
        ur = ul + (sl - sr);
        vr = vl;
        ul = ue;
        vl = ve;
        double a = ur, b = vr;
        ur = a * nx + b * tx;
        vr = a * ny + b * ty;
      }
      int set_wall = 0;
      if (s->o.hvalve) {
        double dot = ur * nx + vr * ny;
        if (dot < ZERO) { set_wall = 1; hr = hl; ur = ZERO; vr = ZERO; }
      }
      double f[3];
      flux_ac(nx, ny, hl, ul, vl, hr, ur, vr, f);
      if (set_wall) f[0] = ZERO;
      const double L = s->len[k];
      s->phi[i] = s->phi[i] - f[0] * L;
      s->phi[s->N + i] = s->phi[s->N + i] - f[1] * L;
      s->phi[2 * s->N + i] = s->phi[2 * s->N + i] - f[2] * L;
    }
  }
}

/* crit_outflow.f90:29-107 (htype 2: Froude number 1) */
static void crit_outflow(orc_t *s) {
  const double h_dry = s->o.h_dry;
  for (int ch = 0; ch < s->n_out; ++ch) {
    if (s->out_type[ch] != 2) continue;
    for (int l = s->out_ptr[ch]; l < s->out_ptr[ch + 1]; ++l) {
      int k = s->out_edges[l];
      int i = edge_in_element(s, k);
      int m = local_edge(s, i, k);
      double hl, ue, ve, tmpu, tmpv;
      face(s, i, m, &hl, &tmpu, &tmpv);
      if (s->dry[i] || hl < h_dry) { ue = ZERO; ve = ZERO; hl = fmax2(hl, ZERO); }
      else { ue = tmpu; ve = tmpv; }
      double nx = s->nrm[2 * k], ny = s->nrm[2 * k + 1], tx = s->tng[2 * k], ty = s->tng[2 * k + 1];
      if (s->gsign[3 * i + m] < 0) { nx = -nx; ny = -ny; tx = -tx; ty = -ty; }
      if (hl < h_dry) continue;
      double ul = ue * nx + ve * ny;
      double vl = ue * tx + ve * ty;
      double ur = ul, vr = vl;
      double hr = ur * ur / G;
      double a = ur, b = vr;
      ur = a * nx + b * tx;
      vr = a * ny + b * ty;
      if (s->o.hvalve) {
        double dot = ur * nx + vr * ny;
        if (dot < ZERO) { hr = hl; ur = ZERO; vr = ZERO; }
      }
      double f[3];
      f[0] = hr * (ur * nx + vr * ny);
      f[1] = hr * ur * (ur * nx + vr * ny) + HALF * G * hr * hr * nx;
      f[2] = hr * vr * (ur * nx + vr * ny) + HALF * G * hr * hr * ny;
      double dh = ZERO;
      const double L = s->len[k];
      s->phi[i] = s->phi[i] - f[0] * L;
      s->phi[s->N + i] = s->phi[s->N + i] - (f[1] + dh * nx) * L;
      s->phi[2 * s->N + i] = s->phi[2 * s->N + i] - (f[2] + dh * ny) * L;
    }
  }
}

/* free_flow.f90:30-93 (htype 0) */
static void free_flow(orc_t *s) {
  const double h_dry = s->o.h_dry;
  for (int ch = 0; ch < s->n_out; ++ch) {
    if (s->out_type[ch] != 0) continue;
    for (int l = s->out_ptr[ch]; l < s->out_ptr[ch + 1]; ++l) {
      int k = s->out_edges[l];
      int i = edge_in_element(s, k);
      int m = local_edge(s, i, k);
      double hl, ue, ve, tmpu, tmpv;
      face(s, i, m, &hl, &tmpu, &tmpv);
      hl = fmax2(hl, ZERO);
      if (hl < h_dry) continue;
      if (s->h[i] < h_dry) { ue = ZERO; ve = ZERO; } else { ue = tmpu; ve = tmpv; }
      double nx = s->nrm[2 * k], ny = s->nrm[2 * k + 1];
      if (s->gsign[3 * i + m] < 0) { nx = -nx; ny = -ny; }
      double ul = ue, vl = ve, hr = hl, ur = ul, vr = vl;
      int set_wall = 0;
      if (s->o.hvalve) {
        double dot = ur * nx + vr * ny;
        if (dot < 0) { set_wall = 1; ur = ZERO; vr = ZERO; }
      }
      double f[3];
      flux_ac(nx, ny, hl, ul, vl, hr, ur, vr, f);
      if (set_wall) f[0] = ZERO;
      for (int n = 0; n < 3; ++n) s->phi[n * s->N + i] = s->phi[n * s->N + i] - f[n] * s->len[k];
    }
  }
}

/* fric_terms.f90:24-98 (explicit Manning friction; a stage source) */
/* fric_terms.f90:30-38,87-96 (Manning, opt_friction 0) on n cells. */
static void fric_explicit(int n, const unsigned char *dry, const double *cd, const double *h,
                          const double *umag, const double *u, const double *v,
                          const double *area, double *phi2, double *phi3) {
  for (int i = 0; i < n; ++i) {
    double r = dry[i] ? ZERO : G * cd[i] * cd[i] / fric_pow(h[i]);
    r = -r * umag[i];
    phi2[i] = phi2[i] + r * u[i] * area[i];
    phi3[i] = phi3[i] + r * v[i] * area[i];
  }
}

static void fric_terms(orc_t *s) {
  fric_explicit(s->N, s->dry, s->cd, s->h, s->u_mag, s->u, s->v, s->area, s->phi + s->N,
                s->phi + 2 * (size_t)s->N);
}

void orc_fric_explicit(int32_t n, const int32_t *dry, const double *cd, const double *h,
                       const double *umag, const double *u, const double *v, const double *area,
                       double *phi2, double *phi3) {
  unsigned char *d = xcalloc(n, 1);
  for (int i = 0; i < n; ++i) d[i] = dry[i] != 0;
  fric_explicit(n, d, cd, h, umag, u, v, area, phi2, phi3);
  free(d);
}

/* The node-based routines of the reference's RDS solver, on this mesh with a
 * bed z and depth h given at the vertices:
 *  - fvsources.f90:27-46: the node least-squares bed gradient
 *    (lsq_gradient.f90:35-47 over n2n, weights 1/d^2 of lsq_setup.f90:32-80,
 *    solver 1, type 2) times -h, integrated over the median dual
 *    (integral_md.f90:33-56; its c1 and c2 are default-real literals, and c1
 *    is 2.037..., not the 11/54 its comment gives);
 *  - drynwet.f90:28-50 with part_dry.f90:30-43 per triangle (zadjust on):
 *    the BGN bed of partially dry triangles.  With one dry vertex i the
 *    reference takes vertices MAX(1,MOD(i+1,4)) and MAX(1,MOD(i+2,4)), which
 *    for i = 3 are both vertex 1.
 * source: [3P] component-major; zout: [3N]; pd: [4N] = part_dry's flag, n,
 * i, j (1-based, 0 where unset). */
void orc_rds(const orc_t *s, double h_dry, const double *z, const double *h, double *source,
             double *zout, int32_t *pd) {
  const int P = s->P, N = s->N, E = s->E;
  /* node_db.f90:165-185: n2n from the edge table */
  int *np = xcalloc(P + 1, sizeof(int)), *nn = xcalloc(2 * (size_t)E + 1, sizeof(int));
  for (int k = 0; k < E; ++k) { np[s->ep[2 * k] + 1]++; np[s->ep[2 * k + 1] + 1]++; }
  for (int i = 0; i < P; ++i) np[i + 1] += np[i];
  {
    int *fill = xcalloc(P, sizeof(int));
    for (int k = 0; k < E; ++k) {
      const int p1 = s->ep[2 * k], p2 = s->ep[2 * k + 1];
      nn[np[p1] + fill[p1]++] = p2;
      nn[np[p2] + fill[p2]++] = p1;
    }
    free(fill);
  }
  double *zbx = xcalloc(P, 8), *zby = xcalloc(P, 8);
  for (int i = 0; i < P; ++i) {
    double ai = ZERO, bi = ZERO, ci = ZERO, di = ZERO, ei = ZERO;
    for (int t = np[i]; t < np[i + 1]; ++t) {
      const int l = nn[t];
      const double dx = s->x[l] - s->x[i], dy = s->y[l] - s->y[i];
      const double w = ONE / (dx * dx + dy * dy);
      ai = ai + w * dx * dx;
      bi = bi + w * dx * dy;
      ci = ci + w * dy * dy;
    }
    const double det = ai * ci - bi * bi;
    for (int t = np[i]; t < np[i + 1]; ++t) {
      const int l = nn[t];
      const double dx = s->x[l] - s->x[i], dy = s->y[l] - s->y[i];
      const double w = ONE / (dx * dx + dy * dy);
      const double dphi = z[l] - z[i];
      di = di + w * dphi * dx;
      ei = ei + w * dphi * dy;
    }
    const double zx = (ci * di - bi * ei) / det, zy = (ai * ei - di * bi) / det;
    zbx[i] = -zx * h[i];
    zby[i] = -zy * h[i];
  }
  const double c1 = (double)2.03703703703704f, c2 = (double)6.481481481481481E-2f;
  for (int i = 0; i < P; ++i) {
    double ix = ZERO, iy = ZERO;
    for (int t = s->n2t_ptr[i]; t < s->n2t_ptr[i + 1]; ++t) {
      const int c = s->n2t[t];
      int m = 0;
      while (s->vtx[3 * c + m] != i) ++m;
      const int j = s->vtx[3 * c + (m + 1) % 3], k = s->vtx[3 * c + (m + 2) % 3];
      ix = ix + s->area[c] * (c1 * zbx[i] + c2 * (zbx[j] + zbx[k]));
      iy = iy + s->area[c] * (c1 * zby[i] + c2 * (zby[j] + zby[k]));
    }
    source[i] = ZERO;
    source[P + i] = G * ix;
    source[2 * (size_t)P + i] = G * iy;
  }
  for (int e = 0; e < N; ++e) {
    const int *v = &s->vtx[3 * e];
    double *zo = &zout[3 * (size_t)e];
    int32_t *q = &pd[4 * (size_t)e];
    for (int m = 0; m < 3; ++m) zo[m] = z[v[m]];
    int n = 0, fi = 0, fj = 0;
    for (int m = 1; m <= 3; ++m)
      if (h[v[m - 1]] < h_dry) { n = n + 1; fj = m; if (n == 1) fi = m; }
    q[0] = (n != 0 && n != 3); q[1] = n; q[2] = fi; q[3] = fj;
    if (!q[0]) continue;
    if (n == 1) {
      int jj = (fi + 1) % 4, kk = (fi + 2) % 4;
      if (jj < 1) jj = 1;
      if (kk < 1) kk = 1;
      const double dj = z[v[jj - 1]] + h[v[jj - 1]], dk = z[v[kk - 1]] + h[v[kk - 1]];
      const double dmax = fmax2(dj, dk);
      zo[fi - 1] = fmin2(dmax, z[v[fi - 1]]);
    } else {
      const int kk = 6 - fi - fj;
      const double dk = z[v[kk - 1]] + h[v[kk - 1]];
      zo[fi - 1] = fmin2(dk, z[v[fi - 1]]);
      zo[fj - 1] = fmin2(dk, z[v[fj - 1]]);
    }
  }
  free(np); free(nn); free(zbx); free(zby);
}

/* bed_slope_STD.f90:23-70 with adjustZb.f90:41-70 */
static void bed_slope_std(orc_t *s) {
  for (int e = 0; e < s->N; ++e) {
    if (s->dry[e]) continue;
    double hp[3], gx, gy;
    int adjust = 0;
    for (int i = 0; i < 3; ++i) {
      int j = s->vtx[3 * e + i];
      hp[i] = s->h[e] + s->delh[2 * e] * (s->x[j] - s->xc[e]) + s->delh[2 * e + 1] * (s->y[j] - s->yc[e]);
      if (hp[i] < ZERO) adjust = 1;
    }
    if (adjust) {
      double zp[3], xx[3], yy[3], b[3], c[3];
      for (int i = 0; i < 3; ++i) {
        int j = s->vtx[3 * e + i];
        hp[i] = fmin2(ZERO, hp[i]);
        zp[i] = s->zvtx[j] + hp[i];
        xx[i] = s->x[j]; yy[i] = s->y[j];
      }
      b[0] = yy[1] - yy[2]; c[0] = xx[2] - xx[1];
      b[1] = yy[2] - yy[0]; c[1] = xx[0] - xx[2];
      b[2] = yy[0] - yy[1]; c[2] = xx[1] - xx[0];
      double a = HALF / s->area[e];
      gx = a * (b[0] * zp[0] + b[1] * zp[1] + b[2] * zp[2]);
      gy = a * (c[0] * zp[0] + c[1] * zp[1] + c[2] * zp[2]);
    } else {
      gx = s->delz[2 * e];
      gy = s->delz[2 * e + 1];
    }
    s->phi[s->N + e] = s->phi[s->N + e] - G * s->h[e] * gx * s->area[e];
    s->phi[2 * s->N + e] = s->phi[2 * s->N + e] - G * s->h[e] * gy * s->area[e];
  }
}

/* bed_slope_VTX.f90:22-38 */
static void bed_slope_vtx(orc_t *s) {
  for (int i = 0; i < s->N; ++i) {
    int a = s->vtx[3 * i], b = s->vtx[3 * i + 1], c = s->vtx[3 * i + 2];
    double x1 = s->x[a], y1 = s->y[a], x2 = s->x[b], y2 = s->y[b], x3 = s->x[c], y3 = s->y[c];
    double z1 = s->zvtx[a], z2 = s->zvtx[b], z3 = s->zvtx[c];
    s->delz[2 * i] = HALF * (z1 * (y2 - y3) + z2 * (y3 - y1) + z3 * (y1 - y2)) / s->area[i];
    s->delz[2 * i + 1] = HALF * (z1 * (x3 - x2) + z2 * (x1 - x3) + z3 * (x2 - x1)) / s->area[i];
  }
}

/* get_time_bcs.f90:31-77 */
static void get_time_bcs(orc_t *s, double time) {
  for (int k = 0; k < s->n_out; ++k) {
    if (s->out_type[k] != 1) continue;
    int a = s->hts_ptr[k], n = s->hts_ptr[k + 1] - a;
    if (n <= 0) continue;
    const double *tl = s->hts_t + a, *tv = s->hts_v + a;
    if (time < tl[0]) s->hbc[k] = tv[0];
    else if (time >= tl[n - 1]) s->hbc[k] = tv[n - 1];
    else
      for (int i = 0; i < n - 1; ++i)
        if (time >= tl[i] && time < tl[i + 1]) {
          double d = time - tl[i], d2 = tl[i + 1] - tl[i];
          s->hbc[k] = tv[i] + d * (tv[i + 1] - tv[i]) / d2;
          break;
        }
  }
  for (int k = 0; k < s->n_in; ++k) {
    int a = s->qts_ptr[k], n = s->qts_ptr[k + 1] - a;
    if (n <= 0) continue;
    const double *tl = s->qts_t + a, *tv = s->qts_v + a;
    if (time < tl[0]) s->qin[k] = tv[0];
    else if (time >= tl[n - 1]) s->qin[k] = tv[n - 1];
    else
      for (int i = 0; i < n - 1; ++i)
        if (time >= tl[i] && time < tl[i + 1]) {
          double d = time - tl[i], d2 = tl[i + 1] - tl[i];
          s->qin[k] = tv[i] + d * (tv[i + 1] - tv[i]) / d2;
          break;
        }
  }
}

/* fric_terms_impl.f90:32-99 (Manning, opt_friction 0) */
static void fric_terms_impl(orc_t *s, double delta_t) {
  const int R = s->o.rkorder;
  for (int i = 0; i < s->N; ++i) s->rough[i] = ZERO;
  for (int i = 0; i < s->N; ++i)
    if (!s->dry[i]) s->rough[i] = G * s->cd[i] * s->cd[i] / fric_pow(s->h[i]);
  for (int i = 0; i < s->N; ++i)
    if (!s->dry[i]) s->rough[i] = s->h[i] / (s->h[i] + delta_t * s->rough[i] * s->u_mag[i]);
  for (int i = 0; i < s->N; ++i) { s->u[i] = s->u[i] * s->rough[i]; s->v[i] = s->v[i] * s->rough[i]; }
  for (int i = 0; i < s->N; ++i)
    if (!s->dry[i]) { RK(s, 1, R, i) = s->h[i] * s->u[i]; RK(s, 2, R, i) = s->h[i] * s->v[i]; }
}

/* The CFL time step in corrected form: deltatFVT.f90:32-38 indexes the
 * node-sized cv_area/cv_perim with cell ids (out of bounds) and reads u_mag
 * and wavec left from the previous step.  Here: the minimum over cells of
 * cfl*A/((|u|+c+small)*P), with |u| and c from find_dry_cells applied to the
 * state at the start of the step. */
static double dt_cfl(orc_t *s) {
  find_dry_cells(s, 1);
  double dt = DBL_MAX;
  for (int i = 0; i < s->N; ++i) {
    double d = s->u_mag[i] + s->wavec[i] + SMALL;
    d = s->o.cfl * s->area[i] / (d * s->perim[i]);
    dt = fmin2(d, dt);
  }
  return dt;
}

double orc_dt_cfl(orc_t *s) { return dt_cfl(s); }

static void start(orc_t *s) {
  /* solverFVT.f90:53-127 */
  s->iter = 0; s->rmax = ZERO; s->stime = ZERO; s->ttime = ZERO; s->delt = ZERO;
  memset(s->delh, 0, sizeof(double) * 2 * s->N);
  memset(s->delu, 0, sizeof(double) * 2 * s->N);
  memset(s->delv, 0, sizeof(double) * 2 * s->N);
  bed_slope_vtx(s);
  get_time_bcs(s, s->stime);
  enforce_bcs_h(s);
  for (int i = 0; i < s->N; ++i) s->rough[i] = s->cd[i];
  const int R = s->o.rkorder;
  for (int i = 0; i < s->N; ++i) {
    RK(s, 1, R, i) = s->h[i] * s->u[i];
    RK(s, 2, R, i) = s->h[i] * s->v[i];
    RK(s, 0, R, i) = s->h[i];
  }
  s->started = 1;
}

/* one time step: solverFVT.f90:132-476.  Returns 1 when the run stops. */
/* windforcing.f90:64-86 (opt_solver 2, 3): cell values as given, vertex
 * values averaged by vtx2ctr; wcoef = wcoef0 w |w| / rho */
int32_t orc_set_wind(orc_t *s, int32_t n, const double *c, const double *wx, const double *wy) {
  const int N = s->N;
  free(s->wcoef1); free(s->wcoef2);
  s->wcoef1 = s->wcoef2 = NULL;
  if (n == 0) return 0;
  if (n != N && n != s->P) return -1;
  double *w0 = xcalloc(N, sizeof(double));
  s->wcoef1 = xcalloc(N, sizeof(double));
  s->wcoef2 = xcalloc(N, sizeof(double));
  if (n == N) {
    memcpy(w0, c, sizeof(double) * N);
    memcpy(s->wcoef1, wx, sizeof(double) * N);
    memcpy(s->wcoef2, wy, sizeof(double) * N);
  } else {
    vtx2ctr(s, c, w0);
    vtx2ctr(s, wx, s->wcoef1);
    vtx2ctr(s, wy, s->wcoef2);
  }
  for (int i = 0; i < N; ++i) {
    double mag = sqrt(s->wcoef1[i] * s->wcoef1[i] + s->wcoef2[i] * s->wcoef2[i]);
    s->wcoef1[i] = w0[i] * s->wcoef1[i] * mag / RHO;
    s->wcoef2[i] = w0[i] * s->wcoef2[i] * mag / RHO;
  }
  free(w0);
  return 0;
}

/* pto_in_triangle.f90:20-43 */
static int pto_in_triangle(const orc_t *s, int e, double x, double y) {
  const double tarea = s->area[e];
  const int a = s->vtx[3 * e], b = s->vtx[3 * e + 1], c = s->vtx[3 * e + 2];
  const double x1 = s->x[a], y1 = s->y[a], x2 = s->x[b], y2 = s->y[b], x3 = s->x[c], y3 = s->y[c];
  const double sv = HALF / tarea * (y1 * x3 - x1 * y3 + (y3 - y1) * x + (x1 - x3) * y);
  const double t = HALF / tarea * (x1 * y2 - y1 * x2 + (y1 - y2) * x + (x2 - x1) * y);
  if (sv < ZERO) return 0;
  if (t < ZERO) return 0;
  if (ONE - sv - t < ZERO) return 0;
  return 1;
}

/* ioculvert.f90's data, located by cvprep.f90:29-58: the first triangle (in
 * cell order) holding each inlet and outlet point, and the inlet's offset
 * from that triangle's centroid */
int32_t orc_set_culverts(orc_t *s, int32_t n, const double *pin, const double *pout,
                         const int32_t *ntab, const double *head, const double *q) {
  void *old[] = {s->cvin, s->cvout, s->cvtab_ptr, s->cvpin, s->cvrin, s->cvhead, s->cvq, s->cvsrc};
  for (size_t k = 0; k < sizeof(old) / sizeof(old[0]); ++k) free(old[k]);
  s->cvin = s->cvout = s->cvtab_ptr = NULL;
  s->cvpin = s->cvrin = s->cvhead = s->cvq = s->cvsrc = NULL;
  s->ncv = 0;
  if (n <= 0) return 0;
  s->cvin = xcalloc(n, sizeof(int));
  s->cvout = xcalloc(n, sizeof(int));
  s->cvtab_ptr = xcalloc(n + 1, sizeof(int));
  s->cvpin = xcalloc(3 * n, sizeof(double));
  s->cvrin = xcalloc(2 * n, sizeof(double));
  s->cvsrc = xcalloc(n, sizeof(double));
  for (int i = 0; i < n; ++i) s->cvtab_ptr[i + 1] = s->cvtab_ptr[i] + ntab[i];
  s->cvhead = xcalloc(s->cvtab_ptr[n] + 1, sizeof(double));
  s->cvq = xcalloc(s->cvtab_ptr[n] + 1, sizeof(double));
  memcpy(s->cvhead, head, sizeof(double) * s->cvtab_ptr[n]);
  memcpy(s->cvq, q, sizeof(double) * s->cvtab_ptr[n]);
  memcpy(s->cvpin, pin, sizeof(double) * 3 * n);
  s->ncv = n;
  for (int i = 0; i < n; ++i) {
    int j, found = -1;
    for (j = 0; j < s->N; ++j)
      if (pto_in_triangle(s, j, pin[3 * i], pin[3 * i + 1])) { found = j; break; }
    if (found < 0) { s->ncv = 0; return i + 1; }
    s->cvin[i] = found;
    s->cvrin[2 * i] = pin[3 * i] - s->xc[found];
    s->cvrin[2 * i + 1] = pin[3 * i + 1] - s->yc[found];
    found = -1;
    for (j = 0; j < s->N; ++j)
      if (pto_in_triangle(s, j, pout[2 * i], pout[2 * i + 1])) { found = j; break; }
    if (found < 0) { s->ncv = 0; return -(i + 1); }
    s->cvout[i] = found;
  }
  return 0;
}

void orc_get_culverts(const orc_t *s, int32_t *cin, int32_t *cout, double *src) {
  for (int i = 0; i < s->ncv; ++i) {
    if (cin) cin[i] = s->cvin[i];
    if (cout) cout[i] = s->cvout[i];
    if (src) src[i] = s->cvsrc[i];
  }
}

/* cvdisch.f90:26-58: the discharge of each culvert from the water surface
 * reconstructed at its inlet (delh + delz is the stage gradient) */
static void cvdisch(orc_t *s) {
  for (int i = 0; i < s->ncv; ++i) {
    s->cvsrc[i] = ZERO;
    const int e = s->cvin[i];
    if (s->dry[e]) continue;
    double head = s->zeta[e] + (s->delh[2 * e] + s->delz[2 * e]) * s->cvrin[2 * i] +
                  (s->delh[2 * e + 1] + s->delz[2 * e + 1]) * s->cvrin[2 * i + 1];
    head = head - s->cvpin[3 * i + 2];
    if (head < s->o.h_dry) continue;
    const double *hh = s->cvhead + s->cvtab_ptr[i], *qq = s->cvq + s->cvtab_ptr[i];
    const int k = s->cvtab_ptr[i + 1] - s->cvtab_ptr[i];
    if (head < hh[0]) {
      s->cvsrc[i] = qq[0];
    } else if (head >= hh[k - 1]) {
      s->cvsrc[i] = qq[k - 1];
    } else {
      for (int j = 0; j < k - 1; ++j)
        if (head < hh[j + 1]) {
          double t = (head - hh[j]) / (hh[j + 1] - hh[j]);
          s->cvsrc[i] = qq[j + 1] * t + qq[j] * (ONE - t);
          break;
        }
    }
  }
}

/* solverFVT.f90:238-256: Gauss-Seidel sweeps of the residual Laplacian in
 * cell order over elemnt_db's t2t; fac1, fac2 of :59-60 */
static void rsavg(orc_t *s) {
  const int N = s->N;
  const double eps = s->o.rsavg_eps;
  const double fac1 = ONE / (ONE + eps * 3.0);
  const double fac2 = eps / (ONE + eps * 3.0);
  if (!s->phi0) s->phi0 = xcalloc(3 * (size_t)N, sizeof(double));
  memcpy(s->phi0, s->phi, sizeof(double) * 3 * N);
  for (int k = 0; k < s->o.rsavg_n; ++k)
    for (int i = 0; i < N; ++i) {
      for (int v = 0; v < 3; ++v) s->phi[v * N + i] = fac1 * s->phi0[v * N + i];
      for (int m = s->et2t_ptr[i]; m < s->et2t_ptr[i + 1]; ++m) {
        const int j = s->et2t[m];
        for (int v = 0; v < 3; ++v) s->phi[v * N + i] = s->phi[v * N + i] + fac2 * s->phi[v * N + j];
      }
    }
}

static int step(orc_t *s) {
  const int N = s->N, R = s->o.rkorder;
  const double h_dry = s->o.h_dry;
  s->iter = s->iter + 1;
  if (s->iter > s->o.istop) { s->iter = s->iter - 1; return 1; }
  for (int v = 0; v < 3; ++v)
    for (int i = 0; i < N; ++i) RK(s, v, 0, i) = RK(s, v, R, i);
  if (!s->o.steady) s->delt = (s->o.dt_op == 0) ? s->o.cfl : dt_cfl(s);
  s->ttime = s->ttime + s->delt;
  for (int rks = 1; rks <= R; ++rks) {
    for (int i = 0; i < N; ++i) s->zeta[i] = s->h[i] + s->z[i];
    find_dry_cells(s, 1);
    cell_grads(s);
    hgradfix2(s, rks);
    to_vertices(s);
    memset(s->phi, 0, sizeof(double) * 3 * N);
    interior_fluxes(s);
    if (s->o.btype == 0) wenslp3(s); else weslp2(s);
    inflow_q(s, 1);
    inflow_q(s, 5);
    inflow_h(s);
    outflow_h(s);
    crit_outflow(s);
    free_flow(s);
    if (s->o.friction == 1) fric_terms(s);
    bed_slope_std(s);
    if (s->wcoef1) {  /* solverFVT.f90:218-224 */
      for (int i = 0; i < N; ++i) {
        if (s->dry[i]) continue;
        s->phi[N + i] = s->phi[N + i] + s->wcoef1[i];
        s->phi[2 * N + i] = s->phi[2 * N + i] + s->wcoef2[i];
      }
    }
    if (s->ncv > 0) {  /* solverFVT.f90:227-233 */
      cvdisch(s);
      for (int i = 0; i < s->ncv; ++i) {
        s->phi[s->cvin[i]] = s->phi[s->cvin[i]] - s->cvsrc[i];
        s->phi[s->cvout[i]] = s->phi[s->cvout[i]] + s->cvsrc[i];
      }
    }
    if (s->o.rsavg) rsavg(s);
    /* solverFVT.f90:259-283 */
    for (int i = 0; i < N; ++i) {
      if (s->o.steady) {
        double d = s->u_mag[i] + s->wavec[i];
        if (d < SMALL) d = 1.044;
        s->delt = s->o.cfl * s->area[i] / (d * s->perim[i]);
      }
      double dt = s->delt / s->area[i];
      for (int v = 0; v < 3; ++v) {
        double acc = ZERO;
        for (int j = 0; j <= rks - 1; ++j) acc = acc + s->rka[rks - 1][j] * RK(s, v, j, i);
        acc = acc + s->rkb[rks - 1] * dt * s->phi[v * N + i];
        RK(s, v, rks, i) = acc;
      }
    }
    if (s->o.vclip) {  /* solverFVT.f90:286-296 */
      for (int i = 0; i < N; ++i) {
        double heff = fmax2(h_dry, RK(s, 0, rks, i));
        double ru = RK(s, 1, rks, i), rv = RK(s, 2, rks, i);
        double um = sqrt(ru * ru + rv * rv) / heff;
        if (um > s->o.vceiling) {
          RK(s, 1, rks, i) = s->o.vceiling * ru / um;
          RK(s, 2, rks, i) = s->o.vceiling * rv / um;
        }
      }
    }
    /* solverFVT.f90:300-325 */
    for (int i = 0; i < N; ++i) {
      RK(s, 0, rks, i) = fmax2(RK(s, 0, rks, i), ZERO);
      s->h[i] = RK(s, 0, rks, i);
      double hd = s->dry[i] ? s->o.h_wet : h_dry;
      if (s->h[i] > hd) {
        s->u[i] = RK(s, 1, rks, i) / s->h[i];
        s->v[i] = RK(s, 2, rks, i) / s->h[i];
      } else {
        RK(s, 1, rks, i) = ZERO; RK(s, 2, rks, i) = ZERO;
        s->u[i] = ZERO; s->v[i] = ZERO;
      }
    }
  }
  find_dry_cells(s, 0);
  if (s->o.friction == 0) fric_terms_impl(s, s->delt);
  double sumtop = ZERO, sumbot = ZERO, sumtopq = ZERO, sumbotq = ZERO;
  for (int i = 0; i < N; ++i) {
    double d = s->h[i] - RK(s, 0, 0, i);
    sumtop = sumtop + d * d;
    sumbot = sumbot + RK(s, 0, 0, i) * RK(s, 0, 0, i);
  }
  if (s->o.stop_cr == 2) {
    /* stopping test 2 (no reference counterpart): the relative RMS change
       of the discharge (hu, hv) after friction, beside that of h */
    for (int i = 0; i < N; ++i) {
      double qu0 = RK(s, 1, 0, i), qv0 = RK(s, 2, 0, i);
      double du = RK(s, 1, R, i) - qu0, dv = RK(s, 2, R, i) - qv0;
      sumtopq = sumtopq + (du * du + dv * dv);
      sumbotq = sumbotq + (qu0 * qu0 + qv0 * qv0);
    }
  }
  if (s->o.hvalve) {  /* solverFVT.f90:357-370 */
    for (int j = 0; j < s->nbc; ++j) {
      if (!s->hbar[j]) continue;
      int k = s->bcedges[j];
      int i = edge_in_element(s, k);
      int m = local_edge(s, i, k);
      double ks = (s->gsign[3 * i + m] < 0) ? -ONE : ONE;
      double dot = ks * (s->u[i] * s->nrm[2 * k] + s->v[i] * s->nrm[2 * k + 1]);
      if (dot < 0) { s->u[i] = ZERO; s->v[i] = ZERO; }
    }
  }
  for (int i = 0; i < N; ++i)
    if (s->dry[i]) { s->phi[i] = ZERO; s->phi[N + i] = ZERO; s->phi[2 * N + i] = ZERO; }
  int stop = 0;
  if (s->o.steady) {
    if (s->o.stop_cr == 0) {
      double r = ZERO;
      for (int i = 0; i < N; ++i)
        r = fmax2(fmax2(fmax2(r, fabs(s->phi[i])), fabs(s->phi[N + i])), fabs(s->phi[2 * N + i]));
      s->rmax = r;
    } else {
      if (orc_equals(sumbot, ZERO)) sumbot = ONE;
      s->rmax = sqrt(sumtop / sumbot);
      if (s->o.stop_cr == 2) {
        if (orc_equals(sumbotq, ZERO)) sumbotq = ONE;
        s->rmax = fmax2(s->rmax, sqrt(sumtopq / sumbotq));
      }
    }
    if (s->rmax < s->o.rstop) stop = 1;
    s->stime = (double)s->iter;
  } else {
    s->stime = s->stime + s->delt;
    double r = ZERO;
    for (int i = 0; i < N; ++i)
      r = fmax2(fmax2(fmax2(r, fabs(s->phi[i])), fabs(s->phi[N + i])), fabs(s->phi[2 * N + i]));
    s->rmax = r;
    get_time_bcs(s, s->stime);
  }
  return stop;
}

int orc_run(orc_t *s, int nsteps, int32_t *iter, double *ttime, double *rmax) {
  if (!s->started) start(s);
  int taken = 0;
  for (int n = 0; n < nsteps; ++n) {
    int before = s->iter;
    int stop = step(s);
    if (s->iter > before) taken++;
    if (stop || s->iter == before) break;
  }
  if (iter) *iter = s->iter;
  if (ttime) *ttime = s->ttime;
  if (rmax) *rmax = s->rmax;
  return taken;
}

/* The post-loop part of solverFVT.f90:483-487: to_vertices (with the vertex
 * velocities and the whole of enforce_bcs), ctr2vtx of cd, error_est(2). */
void orc_finalize(orc_t *s) {
  if (!s->started) start(s);
  to_vertices_final(s);
  ctr2vtx3(s, s->cd, s->cdvtx);
  error_est2(s);
}

void orc_get_output(const orc_t *s, double *uvtx, double *vvtx, double *cdvtx, double *e_est,
                    double *delhx, double *delhy) {
  const int N = s->N, P = s->P;
  if (uvtx) memcpy(uvtx, s->uvtx, sizeof(double) * P);
  if (vvtx) memcpy(vvtx, s->vvtx, sizeof(double) * P);
  if (cdvtx) memcpy(cdvtx, s->cdvtx, sizeof(double) * P);
  if (e_est) memcpy(e_est, s->e_est, sizeof(double) * N);
  for (int i = 0; i < N; ++i) {
    if (delhx) delhx[i] = s->delh[2 * i];
    if (delhy) delhy[i] = s->delh[2 * i + 1];
  }
}

void orc_get_prep(const orc_t *s, double *z, double *zvtx, double *zeta, double *cd,
                  double *cdvtx, double *hvtx, double *uvtx, double *vvtx, double *zetavtx) {
  const int N = s->N, P = s->P;
  if (z) memcpy(z, s->z, sizeof(double) * N);
  if (zvtx) memcpy(zvtx, s->zvtx, sizeof(double) * P);
  if (zeta) memcpy(zeta, s->p_zeta, sizeof(double) * N);
  if (cd) memcpy(cd, s->cd, sizeof(double) * N);
  if (cdvtx) memcpy(cdvtx, s->p_cdvtx, sizeof(double) * P);
  if (hvtx) memcpy(hvtx, s->p_hvtx, sizeof(double) * P);
  if (uvtx) memcpy(uvtx, s->p_uvtx, sizeof(double) * P);
  if (vvtx) memcpy(vvtx, s->p_vvtx, sizeof(double) * P);
  if (zetavtx) memcpy(zetavtx, s->p_zetavtx, sizeof(double) * P);
}

void orc_get_state(const orc_t *s, double *h, double *u, double *v, double *qh, double *qu,
                   double *qv, double *phi, int32_t *dry, double *hvtx, double *zetavtx) {
  const int N = s->N, R = s->o.rkorder;
  if (h) memcpy(h, s->h, sizeof(double) * N);
  if (u) memcpy(u, s->u, sizeof(double) * N);
  if (v) memcpy(v, s->v, sizeof(double) * N);
  for (int i = 0; i < N; ++i) {
    if (qh) qh[i] = RK(s, 0, R, i);
    if (qu) qu[i] = RK(s, 1, R, i);
    if (qv) qv[i] = RK(s, 2, R, i);
    if (dry) dry[i] = s->dry[i];
  }
  if (phi) memcpy(phi, s->phi, sizeof(double) * 3 * N);
  if (hvtx) memcpy(hvtx, s->hvtx, sizeof(double) * s->P);
  if (zetavtx) memcpy(zetavtx, s->zetavtx, sizeof(double) * s->P);
}

/* Every string's table replaced by one constant (as the library's
 * storm_fvt_set_bc_values does; no reference counterpart: a caller of the
 * reference would edit the module tables).  A started run keeps its step and
 * time; the new values replace the current ones at once (htype 1 strings
 * only on the outflow side). */
void orc_set_bc_values(orc_t *s, const double *qin, const double *hbc) {
  if (qin) {
    free(s->qts_t); free(s->qts_v);
    s->qts_t = xcalloc(s->n_in + 1, sizeof(double));
    s->qts_v = xcalloc(s->n_in + 1, sizeof(double));
    for (int k = 0; k <= s->n_in; ++k) s->qts_ptr[k] = k;
    for (int k = 0; k < s->n_in; ++k) {
      s->qts_v[k] = qin[k];
      if (s->started) s->qin[k] = qin[k];
    }
  }
  if (hbc) {
    free(s->hts_t); free(s->hts_v);
    s->hts_t = xcalloc(s->n_out + 1, sizeof(double));
    s->hts_v = xcalloc(s->n_out + 1, sizeof(double));
    for (int k = 0; k <= s->n_out; ++k) s->hts_ptr[k] = k;
    for (int k = 0; k < s->n_out; ++k) {
      s->hts_v[k] = hbc[k];
      if (s->started && s->out_type[k] == 1) s->hbc[k] = hbc[k];
    }
  }
}

void orc_set_state(orc_t *s, const double *h, const double *u, const double *v) {
  memcpy(s->h, h, sizeof(double) * s->N);
  memcpy(s->u, u, sizeof(double) * s->N);
  memcpy(s->v, v, sizeof(double) * s->N);
  s->started = 0;
}

void orc_cell_grads(orc_t *s, double *grad) {
  if (!s->started) start(s);
  for (int i = 0; i < s->N; ++i) s->zeta[i] = s->h[i] + s->z[i];
  find_dry_cells(s, 1);
  cell_grads(s);
  for (int i = 0; i < s->N; ++i) {
    grad[6 * i + 0] = s->delh[2 * i] + s->delz[2 * i];
    grad[6 * i + 1] = s->delh[2 * i + 1] + s->delz[2 * i + 1];
    grad[6 * i + 2] = s->delu[2 * i];
    grad[6 * i + 3] = s->delu[2 * i + 1];
    grad[6 * i + 4] = s->delv[2 * i];
    grad[6 * i + 5] = s->delv[2 * i + 1];
  }
}

int64_t orc_topology(const orc_t *s, int32_t *t2t_ptr, int32_t *t2t, int32_t *t2t3) {
  if (t2t_ptr) memcpy(t2t_ptr, s->t2t_ptr, sizeof(int) * (s->N + 1));
  if (t2t) memcpy(t2t, s->t2t, sizeof(int) * s->t2t_ptr[s->N]);
  if (t2t3) memcpy(t2t3, s->t2t3, sizeof(int) * 3 * s->N);
  return s->t2t_ptr[s->N];
}
