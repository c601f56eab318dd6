/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle of the FVT time step.
 *
 * A plain-C restatement of the reference SToRM cell-centred finite-volume
 * (FVT) path: mesh set-up (find_edges/t_areas/walls/ibc_arrays/node_db/
 * elemnt_db), dataprepFVT, and the solverFVT time loop with every callee on
 * the active path.  Each function cites the reference file:line it follows.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load it, and only as the checker / the timed CPU baseline.  The product
 * path never calls it.
 *
 * All indices are 0-based; -1 means "none".
 */
#ifndef STORM_FVT_ORACLE_H
#define STORM_FVT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_case {
  int32_t n_pts, n_elems, n_edges;
  const double *x, *y;          /* [P] */
  const int32_t *tri;           /* [3N] counter-clockwise vertex ids */
  const int32_t *edge_p;        /* [2E] edges%p */
  const int32_t *edge_e;        /* [2E] edges%e, -1 = none */
  /* datav flags of dataprepFVT.f90:73-140: 1 = value given at vertices */
  int32_t z_at_vtx, h_at_vtx, u_at_vtx, v_at_vtx, cd_at_vtx;
  const double *z, *h, *u, *v, *cd;  /* [P] or [N] per flag */
  /* inflow strings: type (vtype), CSR of node chains, CSR of Q(t) tables */
  int32_t n_in;
  const int32_t *in_type, *in_ptr, *in_nodes, *qts_ptr;
  const double *qts_t, *qts_v;
  /* outflow strings: type (htype), node chains, H(t) tables */
  int32_t n_out;
  const int32_t *out_type, *out_ptr, *out_nodes, *hts_ptr;
  const double *hts_t, *hts_v;
} orc_case_t;

enum { ORC_FLUX_AC = 0, ORC_FLUX_HLL = 1, ORC_FLUX_RUSANOV = 2, ORC_FLUX_BGNVC = 3 };
enum { ORC_GRAD_LSQ = 0, ORC_GRAD_GAUSS = 1, ORC_GRAD_HDLS = 2 };

typedef struct orc_opts {
  int32_t rkorder;     /* 1, 2, 3 (RK ORDER) */
  int32_t dt_op;       /* 0 = fixed dt = cfl, 1 = CFL reduction (corrected deltatFVT) */
  int32_t steady;      /* 1 = steady local time stepping */
  int32_t stop_cr;     /* 0 = max|phi|, 1 = RMS dh, 2 = max(RMS dh, RMS dq) */
  int32_t istop;       /* MAX ITERATION */
  int32_t btype;       /* 0 = no-slip walls (wenslp3), 1 = slip (weslp2) */
  int32_t hvalve;
  int32_t limiter;     /* 0 = zero gradients, 1 = unlimited, 2 = Barth-Jespersen */
  int32_t flux;        /* ORC_FLUX_* on interior edges */
  int32_t gradient;    /* ORC_GRAD_* */
  int32_t friction;    /* 0 = implicit Manning (fric_terms_impl), 1 = explicit (fric_terms) */
  int32_t vclip;
  double cfl, rstop, h_dry, h_wet, vceiling;
  int32_t fluxes;      /* interior-edge loop: 0 = invisc_fluxes.f90, 1 = invisc_fluxes_xsf.f90 */
  int32_t rsavg;       /* ACTIVATE RSAVG: residual smoothing (solverFVT.f90:238-256) */
  int32_t rsavg_n;     /* nRSAVG sweeps */
  double rsavg_eps;    /* epsRSAVG */
} orc_opts_t;

typedef struct orc orc_t;

orc_t *orc_create(const orc_case_t *c, const orc_opts_t *o, char *err, int errlen);
void orc_destroy(orc_t *s);
/* Test probe: move the friction's h**(1/3) by one ulp with probability p
 * per call (p = 0, the default, leaves the reference arithmetic). */
void orc_set_pow_perturbation(double p, uint64_t seed);
/* Advance up to nsteps time steps (solverFVT.f90:129-478); stops early when a
 * steady run converges or iter reaches istop.  Returns steps taken. */
int orc_run(orc_t *s, int nsteps, int32_t *iter, double *ttime, double *rmax);
/* The post-loop part of solverFVT.f90:483-487: to_vertices (which zeroes
 * cell velocities in partially wet cells and forms the vertex fields, the
 * velocities under enforce_bcs included), ctr2vtx(cd), error_est(2). */
void orc_finalize(orc_t *s);
/* After orc_finalize: uvtx, vvtx, cdvtx [P]; e_est, delh (x, y) [N]. */
void orc_get_output(const orc_t *s, double *uvtx, double *vvtx, double *cdvtx, double *e_est,
                    double *delhx, double *delhy);
/* The module copies dataprepFVT.f90:69-140 leaves: z, zeta, cd [N];
 * zvtx, cdvtx, hvtx, uvtx, vvtx, zetavtx [P]. */
void orc_get_prep(const orc_t *s, double *z, double *zvtx, double *zeta, double *cd,
                  double *cdvtx, double *hvtx, double *uvtx, double *vvtx, double *zetavtx);
/* h,u,v: cell state; qh,qu,qv: RK slot rkorder; phi: [3N] component-major;
 * dry: [N]; hvtx,zetavtx: [P].  Any pointer may be NULL. */
void orc_get_state(const orc_t *s, double *h, double *u, double *v, double *qh,
                   double *qu, double *qv, double *phi, int32_t *dry,
                   double *hvtx, double *zetavtx);
void orc_set_state(orc_t *s, const double *h, const double *u, const double *v);
/* Replace every string's table by one constant; a started run goes on with
 * the new values from its next step (qin [n_in], hbc [n_out]; NULL keeps a
 * side). */
void orc_set_bc_values(orc_t *s, const double *qin, const double *hbc);
/* Wind forcing (windforcing.f90:64-86, FVT case): n == N values per cell or
 * n == P per vertex (averaged to the cells by vtx2ctr); wcoef = c w |w| / rho.
 * n = 0 turns it off.  Returns 0, or -1 for another n. */
int32_t orc_set_wind(orc_t *s, int32_t n, const double *c, const double *wx, const double *wy);
/* Culverts (ioculvert.f90's data, cvprep.f90's location): pin [3n] inlet x, y,
 * invert z; pout [2n] outlet x, y; ntab [n]; head, q: the curves end to end.
 * Returns 0, or the 1-based number of a culvert whose inlet (negative: outlet)
 * lies in no triangle. */
int32_t orc_set_culverts(orc_t *s, int32_t n, const double *pin, const double *pout,
                         const int32_t *ntab, const double *head, const double *q);
/* cvtrigin, cvtrigout [n] (0-based cells) and the last stage's cvsrc [n]. */
void orc_get_culverts(const orc_t *s, int32_t *cin, int32_t *cout, double *src);
/* Per-routine entry points (pinned against closed forms). */
void orc_flux(int32_t kind, double nx, double ny, double hl, double ul, double vl,
              double hr, double ur, double vr, double cl, double cr, double h0,
              double *f);
int32_t orc_equals(double a, double b);
/* Cell gradients of the current state after cell_grads (+hgradfix2 off):
 * grad: [6N] = (dzeta/dx, dzeta/dy, du/dx, du/dy, dv/dx, dv/dy) per cell. */
void orc_cell_grads(orc_t *s, double *grad);
double orc_dt_cfl(orc_t *s);
/* Explicit Manning friction (fric_terms.f90:30-38,87-96) on n cells: phi2,
 * phi3 are updated in place; the oracle's step runs the same code. */
void orc_fric_explicit(int32_t n, const int32_t *dry, const double *cd, const double *h,
                       const double *umag, const double *u, const double *v, const double *area,
                       double *phi2, double *phi3);
/* The RDS node routines on this mesh (fvsources; part_dry and drynwet per
 * triangle) with bed z and depth h at the vertices; see fvt_oracle.c. */
void orc_rds(const orc_t *s, double h_dry, const double *z, const double *h, double *source,
             double *zout, int32_t *pd);
int64_t orc_topology(const orc_t *s, int32_t *t2t_ptr, int32_t *t2t, int32_t *t2t3);

#ifdef __cplusplus
}
#endif
#endif
