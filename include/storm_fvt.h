/*
 * storm_fvt.h -- C ABI of the MI355X-native SToRM FVT time step.
 *
 * The library replaces the reference's FVT data preparation and time loop:
 *
 *   storm_fvt_create          Source_xsy-64bit/dataprepFVT.f90:1-342 (called
 *                             at solver.f90:50), plus the mesh tables the
 *                             reference builds in find_edges/t_areas/walls/
 *                             ibc_arrays/node_db/elemnt_db (read_mfa.f90:161-173)
 *   storm_fvt_set_state       the initial h,u,v of dataprepFVT.f90:90-130
 *   storm_fvt_set_bc_values   the Q / stage values of get_time_bcs.f90:31-77
 *   storm_fvt_run             the time loop solverFVT.f90:129-478 (called at
 *                             solver.f90:57), nsteps at a time
 *   storm_fvt_finalize        the post-loop work of solverFVT.f90:483-487:
 *                             to_vertices (to_vertices.f90:1-68 with the whole
 *                             of enforce_bcs.f90), ctr2vtx of cd, error_est(2)
 *   storm_fvt_get_prep_fields the cell and vertex copies dataprepFVT.f90:69-140
 *                             leaves in the modules (z, zvtx, zeta, cd, cdvtx)
 *   storm_fvt_get_state,
 *   storm_fvt_get_conserved,
 *   storm_fvt_get_vertex      what solverFVT leaves in the dep_vars / RKparams
 *                             modules (h,u,v,phi,drycell,rk*(:,rkorder),hvtx,
 *                             uvtx,vvtx,zetavtx; dep_vars.f90:16-57,
 *                             RKparams.f90:19)
 *   storm_fvt_get_output      cdvtx and e_est after the loop (solverFVT.f90:
 *                             484-487, error_est.f90:57-84)
 *   storm_fvt_get_c2v_weights the centre-to-vertex weights of c2v_setup.f90:
 *                             98-119 (type 3), which the module keeps for
 *                             later ctr2vtx calls (write_cgns.f90:182,211)
 *   storm_fvt_get_gradients   delh, the depth gradient of the last stage
 *                             (cell_grads.f90:24-75, hgradfix2.f90:46-98),
 *                             which error_est(2) reads at solverFVT.f90:487
 *   storm_fvt_get_bc_values   qin / hbc as get_time_bcs last left them
 *                             (solverFVT.f90:462), which the post-loop
 *                             enforce_bcs reads (to_vertices.f90:65)
 *   storm_fvt_last_error,
 *   storm_fvt_destroy         the error/shutdown role of byebye.f90:29-215
 *
 * Conventions: plain pointers and sizes, 0-based indices, -1 = none, FP64.
 * Every int-returning call returns 0 on success or a negative STORM_E_* code
 * with a message from storm_fvt_last_error(); no C++ exception crosses the
 * ABI.  Host arrays stay owned by the caller (copied in and out); the context
 * owns its device buffers and one HIP stream.  One context per host thread.
 */
#ifndef STORM_FVT_H
#define STORM_FVT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STORM_FVT_ABI_VERSION 1

enum {
  STORM_OK = 0,
  STORM_E_ARG = -1,       /* invalid argument / option */
  STORM_E_MESH = -2,      /* inconsistent mesh connectivity or boundary data */
  STORM_E_DEVICE = -3,    /* HIP runtime error or no device */
  STORM_E_NOMEM = -4,     /* allocation failure */
  STORM_E_STATE = -5      /* call out of sequence */
};

/* Interior-edge numerical flux (the call at invisc_fluxes.f90:137 and the
 * commented alternatives at :138,143-145). */
enum { STORM_FLUX_AC = 0, STORM_FLUX_HLL = 1, STORM_FLUX_RUSANOV = 2, STORM_FLUX_BGNVC = 3 };
/* Cell gradients (cell_grads.f90:46-54). */
/* LSQ: lsq_grad_array on t2t (weights 1/d^2); GAUSS: Green-Gauss (A&C eqs
 * 11-13); HDLS: hdlsq_grad_array on t2tHD, unit weights (ACTIVATE HDLS,
 * cell_grads.f90:37-42, hdlsq_setup.f90, elemnt_db.f90:136-214). */
enum { STORM_GRAD_LSQ = 0, STORM_GRAD_GAUSS = 1, STORM_GRAD_HDLS = 2 };

/* Mesh and initial data: the contents of the reference's geometry/dep_vars/
 * vbc_arrays modules after read_mfa (tecplot_read.f90, ioflow.f90). */
typedef struct storm_mesh {
  int32_t n_pts, n_elems, n_edges;   /* n_edges = 0: edges are built here
                                        (find_edges.f90:47-54 order) */
  const double *x, *y;               /* [n_pts] node coordinates */
  const int32_t *tri;                /* [3*n_elems] counter-clockwise vertex ids */
  const int32_t *edge_p;             /* [2*n_edges] edges%p (or NULL) */
  const int32_t *edge_e;             /* [2*n_edges] edges%e, -1 = none (or NULL) */
  /* datav flags (dataprepFVT.f90:73-140): 1 = given at vertices, 0 = cells */
  int32_t z_at_vtx, h_at_vtx, u_at_vtx, v_at_vtx, cd_at_vtx;
  const double *z, *h, *u, *v, *cd;  /* bed, depth, velocity, Manning n */
  /* inflow strings (ioflow.f90:184-253): vtype, node chains, Q(t) tables */
  int32_t n_in;
  const int32_t *in_type, *in_ptr, *in_nodes, *qts_ptr;
  const double *qts_t, *qts_v;
  /* outflow strings (ioflow.f90:255-312): htype, node chains, H(t) tables */
  int32_t n_out;
  const int32_t *out_type, *out_ptr, *out_nodes, *hts_ptr;
  const double *hts_t, *hts_v;
} storm_mesh_t;

/* Options (options.f90; defaults of initialize_options.f90:23-141). */
typedef struct storm_fvt_opts {
  int32_t rkorder;   /* RK ORDER: 1, 2, 3 */
  int32_t dt_op;     /* DELT TYPE: 0 fixed dt = cfl; 1 CFL reduction (the
                        corrected deltatFVT: cell area and perimeter) */
  int32_t steady;    /* TIME STEPPING 0: steady, local time steps */
  int32_t stop_cr;   /* STOPPING CRIT: 0 max|phi|, 1 RMS change of h (solverFVT.f90:407-431);
                        2 the larger of the RMS change of h and that of the
                        discharge (hu, hv), each relative as in 1 (no reference
                        counterpart: 1 alone stops runs whose discharge is
                        still far from steady) */
  int32_t istop;     /* MAX ITERATION */
  int32_t btype;     /* WALLS: 0 no-slip (wenslp3), 1 slip (weslp2) */
  int32_t hvalve;    /* HVALVE */
  int32_t limiter;   /* LIMITER: 0 zero gradients, 1 unlimited, 2 Barth-Jespersen */
  int32_t flux;      /* STORM_FLUX_* */
  int32_t gradient;  /* STORM_GRAD_* */
  int32_t friction;  /* 0 implicit Manning (fric_terms_impl), 1 explicit (fric_terms) */
  int32_t vclip;     /* MAX VELOCITY given */
  double cfl, rstop, h_dry, h_wet, vceiling;
  /* MI355X execution options (no reference counterpart) */
  int32_t reorder;   /* cell order on the device: 1 = Hilbert curve (default), 2 = reverse
                        Cuthill-McKee, 0 = input order (results are the same) */
  int32_t block;     /* stage-kernel tile capacity (threads per block): 256, 384, 512; 0 = default */
  int32_t graph;     /* 1 = replay captured HIP graphs of whole steps */
  int32_t nonfinite_stop; /* 1 = a steady run (or reach of a batch) whose residual is
                             not finite stops there, with rmax NaN; 0 = it runs on to
                             MAX ITERATION, as the reference does */
  int32_t fluxes;    /* interior-edge loop: 0 = invisc_fluxes.f90 (the one the
                        reference builds); 1 = invisc_fluxes_xsf.f90 (a dry side
                        keeps its cell depth, the left state takes the right
                        cell's y-gradients of u and v as written at :65-66, an
                        ALGORITHM I1 wall zeroes only the mass flux of flux_ac;
                        needs flux AC and tile capacity 256) */
  int32_t rsavg;     /* ACTIVATE RSAVG: residual smoothing after the source
                        terms (solverFVT.f90:238-256): rsavg_n Gauss-Seidel
                        sweeps in input cell order over elemnt_db's t2t with
                        the factors of :59-60; needs tile capacity 256 */
  int32_t rsavg_n;   /* nRSAVG (reference default 2) */
  double rsavg_eps;  /* epsRSAVG (reference default 0.5), taken as
                        min(|eps|, 1) as read_options.f90:351 does */
} storm_fvt_opts_t;

typedef struct storm_fvt storm_fvt_t;

/* Fills *o with the reference defaults. */
void storm_fvt_default_opts(storm_fvt_opts_t *o);

int storm_fvt_create(const storm_mesh_t *mesh, const storm_fvt_opts_t *opts, storm_fvt_t **ctx);
/* An ensemble of n independent reaches advanced together: every launch of a
 * step covers all of them, and each reach keeps its own time step, time,
 * residual and convergence (a converged steady reach stops while the others
 * go on).  No reference counterpart: the reference runs one reach per
 * process.  The arrays of every other entry point are the reaches' arrays
 * laid end to end in reach order (cells, vertices, strings); iter/ttime/rmax
 * of storm_fvt_run are reach 0's.  Results equal those of n separate
 * contexts.  Steady batches take an even rkorder (STORM_E_ARG otherwise). */
int storm_fvt_create_batch(int32_t n, const storm_mesh_t *meshes, const storm_fvt_opts_t *opts,
                           storm_fvt_t **ctx);
/* Per reach (arrays of n; any may be NULL): steps taken, time, residual of
 * the last step, converged flag. */
int storm_fvt_reach_status(storm_fvt_t *ctx, int32_t *iter, double *ttime, double *rmax,
                           uint8_t *done);
/* Cell-centred h, u, v; restarts the time loop (iter = 0). */
int storm_fvt_set_state(storm_fvt_t *ctx, const double *h, const double *u, const double *v);
/* Replace every string's table by one constant: qin[n_in] (Q, or stage for
 * vtype 4) and hbc[n_out] (stage of htype 1 strings).  NULL keeps a side.
 * A run in progress goes on from its current step and time: the new values
 * replace the current ones at once (those of htype 1 strings only on the
 * outflow side) and get_time_bcs keeps them. */
int storm_fvt_set_bc_values(storm_fvt_t *ctx, const double *qin, const double *hbc);
/* Advance up to nsteps time steps (fewer when istop is reached or a steady
 * run converges).  iter/ttime/rmax as solverFVT's dummy arguments. */
int storm_fvt_run(storm_fvt_t *ctx, int32_t nsteps, int32_t *iter, double *ttime, double *rmax);
/* The work that ends solverFVT (solverFVT.f90:483-487): to_vertices (zeroes
 * u,v in partially wet cells; forms hvtx, zetavtx and the vertex velocities
 * under enforce_bcs), cdvtx = ctr2vtx(cd) and e_est = error_est(2). */
int storm_fvt_finalize(storm_fvt_t *ctx);
/* What dataprepFVT leaves in the modules, in input order (any pointer may be
 * NULL): the cell bed z (vtx2ctr of the vertex bed), zeta, cd [n_elems];
 * zvtx, cdvtx [n_pts] (dataprepFVT.f90:73-140; ctr2vtx type 3 where a field
 * is given at cells, vtx2ctr where at vertices).  Host data: any time. */
int storm_fvt_get_prep_fields(storm_fvt_t *ctx, double *z, double *zvtx, double *zeta, double *cd,
                              double *cdvtx);
/* Cell state in input order; any pointer may be NULL.  phi: [3*n_elems],
 * component-major (phi(:,1), phi(:,2), phi(:,3)). */
int storm_fvt_get_state(storm_fvt_t *ctx, double *h, double *u, double *v, double *phi,
                        uint8_t *drycell);
/* rkh, rku, rkv(:,rkorder): h, hu, hv after the last step. */
int storm_fvt_get_conserved(storm_fvt_t *ctx, double *qh, double *qu, double *qv);
/* The vertex fields [n_pts] of the last to_vertices; any pointer may be
 * NULL.  Before any step or set_state: dataprepFVT's (dataprepFVT.f90:
 * 90-130).  After storm_fvt_finalize: solverFVT's final ones.  After steps
 * with no finalize since: hvtx of the last stage's vertex pass, and
 * STORM_E_STATE for the other three (after set_state with no step since:
 * for all four). */
int storm_fvt_get_vertex(storm_fvt_t *ctx, double *hvtx, double *uvtx, double *vvtx,
                         double *zetavtx);
/* After storm_fvt_finalize (STORM_E_STATE otherwise): cdvtx [n_pts] =
 * ctr2vtx(cd) (solverFVT.f90:484) and e_est [n_elems] (error_est.f90:57-84,
 * etype 2); either may be NULL. */
int storm_fvt_get_output(storm_fvt_t *ctx, double *cdvtx, double *e_est);
/* The centre-to-vertex weights of c2v_setup.f90:98-119 (type 3) as CSR in
 * input numbering: vertex p's cells (node_db.f90:60-68 order) are
 * cells[ptr[p] .. ptr[p+1]), with the normalised weights w (c2v_weights)
 * and the inverse distances wid (c2v_weights_id).  ptr: [n_pts+1].  Returns
 * the number of entries (all pointers NULL: only that) or a negative code. */
int64_t storm_fvt_get_c2v_weights(storm_fvt_t *ctx, int32_t *ptr, int32_t *cells, double *w,
                                  double *wid);
/* delh(:)%x, delh(:)%y [n_elems] in input order, as the last stage formed
 * them from its input state (zero before the first step); either may be
 * NULL. */
int storm_fvt_get_gradients(storm_fvt_t *ctx, double *dhx, double *dhy);
/* storm_fvt_get_gradients without allocating: no reference counterpart;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, hosts that forbid it).  Its staging buffer
 * is a file-scope device array of STORM_FVT_NOALLOC_MAX_CELLS cells (32 MiB
 * per device the library runs on, held until the process ends), shared by
 * every call on a device: calls that overlap on one device race.  Contexts
 * of more cells get STORM_E_ARG; a context with no run started
 * STORM_E_STATE. */
#define STORM_FVT_NOALLOC_MAX_CELLS 2097152
int storm_fvt_get_gradients_noalloc(storm_fvt_t *ctx, double *dhx, double *dhy);
/* The node-based wet/dry and source routines of the reference's RDS solver
 * (not on its FVT path), on the context's mesh for a bed z and depth h given
 * at the vertices [n_pts] (input order; the context's state is not used):
 *   storm_fvt_drynwet   drynwet.f90:28-50 with part_dry.f90:30-43 per
 *                       triangle (zadjust on): zout [3*n_elems], the bed of
 *                       each triangle's vertices (tri order), lowered at the
 *                       dry vertices of a partially dry triangle; pd
 *                       [4*n_elems] or NULL: part_dry's flag, n, i, j
 *                       (1-based, 0 where unset).  h_dry is the context's.
 *   storm_fvt_fvsources fvsources.f90:27-46: source [3*n_pts], component-
 *                       major: 0, then g times the median-dual integral of
 *                       -h times the node least-squares bed gradient (the
 *                       n2n molecule of node_db.f90:165-185, weights of
 *                       lsq_setup.f90:32-80 type 2; integral_md.f90's
 *                       coefficients as written). */
int storm_fvt_drynwet(storm_fvt_t *ctx, const double *z, const double *h, double *zout,
                      int32_t *pd);
int storm_fvt_fvsources(storm_fvt_t *ctx, const double *z, const double *h, double *source);
/* Wind forcing terms (replaces dataprepFVT.f90:237-243 -> windforcing.f90:
 * 43-86, FVT case; applied at solverFVT.f90:218-224): n == n_elems values
 * per cell or n == n_pts per vertex (input order; vertex values are averaged
 * to the cells by vtx2ctr.f90:32-35) of the coefficient c and the wind
 * components wx, wy.  Each wet cell's momentum residuals gain
 * c w |w| / rho (rho = 998.6, constants.f90:29), with no area factor, as the
 * reference adds them.  n = 0 turns the terms off.  Takes effect from the
 * next step. */
int storm_fvt_set_wind(storm_fvt_t *ctx, int64_t n, const double *c, const double *wx,
                       const double *wy);
/* Culverts (replaces dataprepFVT.f90:309 -> cvprep.f90:29-58 on the data
 * ioculvert.f90:84-131 reads; discharges by cvdisch.f90:26-58 every stage,
 * applied at solverFVT.f90:227-233).  n culverts: pin [3n] inlet x, y and
 * invert elevation; pout [2n] outlet x, y; ntab [n] (>= 1) points of each
 * performance curve; head, q: the curves end to end (ascending heads above
 * the invert, discharges).  Each point is placed in the first cell (input
 * order) that holds it; STORM_E_ARG names a point in no cell.  n = 0 removes
 * them.  One reach only (a batch of several takes STORM_E_ARG). */
int storm_fvt_set_culverts(storm_fvt_t *ctx, int32_t n, const double *pin, const double *pout,
                           const int32_t *ntab, const double *head, const double *q);
/* The culverts' cells (cvtrigin, cvtrigout: 0-based input cell ids) and the
 * discharges cvsrc of the latest stage run [n]; any pointer may be NULL. */
int storm_fvt_get_culverts(storm_fvt_t *ctx, int32_t *cell_in, int32_t *cell_out, double *src);
/* The boundary values of the latest step: qin[n_in] (Q, or stage for vtype
 * 4) and hbc[n_out] (stage; meaningful for htype 1 strings).  NULL skips a
 * side. */
int storm_fvt_get_bc_values(storm_fvt_t *ctx, double *qin, double *hbc);
/* The cell count, the library's internal (Hilbert) cell order as the
 * permutation perm[internal] = input id, and the context's hipStream_t, for
 * callers that order their own GPU work after the time loop. */
int storm_fvt_device_info(storm_fvt_t *ctx, int32_t *n_elems, int32_t *perm, void **stream);
/* Seconds of device time of the last storm_fvt_run (HIP events on the
 * context's stream). */
double storm_fvt_last_run_seconds(storm_fvt_t *ctx);
const char *storm_fvt_last_error(storm_fvt_t *ctx);
void storm_fvt_destroy(storm_fvt_t *ctx);

/* Host-only inspection of the set-up (no device needed): the least-squares
 * molecule of every cell (elemnt_db.f90:42-128) as CSR in input numbering.
 * Returns the number of molecule entries or a negative error code. */
int64_t storm_fvt_inspect_molecule(const storm_mesh_t *mesh, int32_t *ptr, int32_t *cells);

#ifdef __cplusplus
}
#endif
#endif
