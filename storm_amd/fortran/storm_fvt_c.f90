! ISO_C_BINDING interface to libstorm_fvt (include/storm_fvt.h).
!
! The derived types mirror storm_mesh_t and storm_fvt_opts_t field by field;
! every function returns 0 or a negative STORM_E_* code.
MODULE storm_fvt_c
  USE, INTRINSIC :: ISO_C_BINDING
  IMPLICIT NONE

  INTEGER(C_INT), PARAMETER :: STORM_OK = 0

  TYPE, BIND(C) :: storm_mesh_t
    INTEGER(C_INT32_T) :: n_pts, n_elems, n_edges
    TYPE(C_PTR) :: x, y, tri, edge_p, edge_e
    INTEGER(C_INT32_T) :: z_at_vtx, h_at_vtx, u_at_vtx, v_at_vtx, cd_at_vtx
    TYPE(C_PTR) :: z, h, u, v, cd
    INTEGER(C_INT32_T) :: n_in
    TYPE(C_PTR) :: in_type, in_ptr, in_nodes, qts_ptr, qts_t, qts_v
    INTEGER(C_INT32_T) :: n_out
    TYPE(C_PTR) :: out_type, out_ptr, out_nodes, hts_ptr, hts_t, hts_v
  END TYPE storm_mesh_t

  TYPE, BIND(C) :: storm_fvt_opts_t
    INTEGER(C_INT32_T) :: rkorder, dt_op, steady, stop_cr, istop, btype, hvalve, &
                          limiter, flux, gradient, friction, vclip
    REAL(C_DOUBLE) :: cfl, rstop, h_dry, h_wet, vceiling
    INTEGER(C_INT32_T) :: reorder, block, graph, nonfinite_stop, fluxes
    INTEGER(C_INT32_T) :: rsavg, rsavg_n
    REAL(C_DOUBLE) :: rsavg_eps
  END TYPE storm_fvt_opts_t

  INTERFACE
    SUBROUTINE storm_fvt_default_opts(o) BIND(C, name="storm_fvt_default_opts")
      IMPORT :: storm_fvt_opts_t
      TYPE(storm_fvt_opts_t), INTENT(OUT) :: o
    END SUBROUTINE

    INTEGER(C_INT) FUNCTION storm_fvt_create(mesh, opts, ctx) BIND(C, name="storm_fvt_create")
      IMPORT :: C_INT, C_PTR, storm_mesh_t, storm_fvt_opts_t
      TYPE(storm_mesh_t), INTENT(IN) :: mesh
      TYPE(storm_fvt_opts_t), INTENT(IN) :: opts
      TYPE(C_PTR), INTENT(OUT) :: ctx
    END FUNCTION

    INTEGER(C_INT) FUNCTION storm_fvt_run(ctx, nsteps, iter, ttime, rmax) BIND(C, name="storm_fvt_run")
      IMPORT :: C_INT, C_PTR, C_INT32_T, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: nsteps
      INTEGER(C_INT32_T), INTENT(OUT) :: iter
      REAL(C_DOUBLE), INTENT(OUT) :: ttime, rmax
    END FUNCTION

    INTEGER(C_INT) FUNCTION storm_fvt_finalize(ctx) BIND(C, name="storm_fvt_finalize")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION

    INTEGER(C_INT) FUNCTION storm_fvt_get_state(ctx, h, u, v, phi, dry) BIND(C, name="storm_fvt_get_state")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, h, u, v, phi, dry
    END FUNCTION

    INTEGER(C_INT) FUNCTION storm_fvt_get_conserved(ctx, qh, qu, qv) BIND(C, name="storm_fvt_get_conserved")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, qh, qu, qv
    END FUNCTION

    INTEGER(C_INT) FUNCTION storm_fvt_get_vertex(ctx, hvtx, uvtx, vvtx, zetavtx) &
        BIND(C, name="storm_fvt_get_vertex")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, hvtx, uvtx, vvtx, zetavtx
    END FUNCTION

    INTEGER(C_INT) FUNCTION storm_fvt_get_prep_fields(ctx, z, zvtx, zeta, cd, cdvtx) &
        BIND(C, name="storm_fvt_get_prep_fields")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, z, zvtx, zeta, cd, cdvtx
    END FUNCTION

    INTEGER(C_INT) FUNCTION storm_fvt_get_output(ctx, cdvtx, e_est) BIND(C, name="storm_fvt_get_output")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, cdvtx, e_est
    END FUNCTION

    INTEGER(C_INT64_T) FUNCTION storm_fvt_get_c2v_weights(ctx, ptr, cells, w, wid) &
        BIND(C, name="storm_fvt_get_c2v_weights")
      IMPORT :: C_INT64_T, C_PTR
      TYPE(C_PTR), VALUE :: ctx, ptr, cells, w, wid
    END FUNCTION

    INTEGER(C_INT) FUNCTION storm_fvt_get_gradients(ctx, dhx, dhy) BIND(C, name="storm_fvt_get_gradients")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, dhx, dhy
    END FUNCTION

    INTEGER(C_INT) FUNCTION storm_fvt_get_bc_values(ctx, qin, hbc) BIND(C, name="storm_fvt_get_bc_values")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, qin, hbc
    END FUNCTION

    FUNCTION storm_fvt_last_error(ctx) BIND(C, name="storm_fvt_last_error")
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: ctx
      TYPE(C_PTR) :: storm_fvt_last_error
    END FUNCTION

    SUBROUTINE storm_fvt_destroy(ctx) BIND(C, name="storm_fvt_destroy")
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END SUBROUTINE

    INTEGER(C_INT) FUNCTION storm_fvt_set_wind(ctx, n, c, wx, wy) BIND(C, name="storm_fvt_set_wind")
      IMPORT :: C_INT, C_PTR, C_INT64_T
      TYPE(C_PTR), VALUE :: ctx, c, wx, wy
      INTEGER(C_INT64_T), VALUE :: n
    END FUNCTION

    INTEGER(C_INT) FUNCTION storm_fvt_set_culverts(ctx, n, pin, pout, ntab, head, q) &
        BIND(C, name="storm_fvt_set_culverts")
      IMPORT :: C_INT, C_PTR, C_INT32_T
      TYPE(C_PTR), VALUE :: ctx, pin, pout, ntab, head, q
      INTEGER(C_INT32_T), VALUE :: n
    END FUNCTION

    INTEGER(C_INT) FUNCTION storm_fvt_get_culverts(ctx, cell_in, cell_out, src) &
        BIND(C, name="storm_fvt_get_culverts")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, cell_in, cell_out, src
    END FUNCTION
  END INTERFACE

CONTAINS

  ! The message of the context's last error as a Fortran string.
  FUNCTION storm_fvt_message(ctx) RESULT(msg)
    TYPE(C_PTR), INTENT(IN) :: ctx
    CHARACTER(LEN=256) :: msg
    CHARACTER(KIND=C_CHAR), POINTER :: cs(:)
    TYPE(C_PTR) :: p
    INTEGER :: i
    msg = ''
    p = storm_fvt_last_error(ctx)
    IF (.NOT. C_ASSOCIATED(p)) RETURN
    CALL C_F_POINTER(p, cs, [256])
    DO i = 1, 256
      IF (cs(i) == C_NULL_CHAR) EXIT
      msg(i:i) = cs(i)
    END DO
  END FUNCTION storm_fvt_message

END MODULE storm_fvt_c
