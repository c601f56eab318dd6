! Drop-in replacements of the reference's FVT entry points,
!   SUBROUTINE dataprepFVT                         (dataprepFVT.f90:1)
!   SUBROUTINE solverFVT(iter,ttime,rmax,fn,title)  (solverFVT.f90:1,35-37)
! with the same names and dummy arguments, called as solver.f90:50,57 calls
! them.  They read the mesh, boundary strings, initial data and options from
! the reference's modules (geometry, dep_vars, options, vbc_arrays), run the
! time step on the GPU through libstorm_fvt (ISO_C_BINDING, storm_fvt_c), and
! leave h, u, v, phi, drycell, rk*(:,rkorder), the vertex fields (hvtx,
! zetavtx, uvtx, vvtx, cdvtx), delh and e_est in the modules where the
! reference's output routines expect them.  A non-zero status of
! the library ends the run through byebye, the reference's fatal-stop path.
MODULE storm_fvt_dropin_state
  USE, INTRINSIC :: ISO_C_BINDING
  IMPLICIT NONE
  SAVE
  TYPE(C_PTR) :: ctx = C_NULL_PTR
END MODULE storm_fvt_dropin_state


SUBROUTINE dataprepFVT
  USE, INTRINSIC :: ISO_C_BINDING
  USE parameters
  USE geometry
  USE dep_vars
  USE options
  USE constants
  USE vbc_arrays
  USE RKparams
  USE storm_fvt_c
  USE storm_fvt_dropin_state
  IMPLICIT NONE

  TYPE(storm_mesh_t) :: m
  TYPE(storm_fvt_opts_t) :: o
  REAL(C_DOUBLE), ALLOCATABLE, TARGET :: xs(:), ys(:), zs(:), hs(:), us(:), vs(:), cds(:)
  INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: tri(:), ep(:), ee(:)
  INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: in_type(:), in_ptr(:), in_nodes(:), qts_ptr(:)
  INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: out_type(:), out_ptr(:), out_nodes(:), hts_ptr(:)
  REAL(C_DOUBLE), ALLOCATABLE, TARGET :: qts_t(:), qts_v(:), hts_t(:), hts_v(:)
  INTEGER :: i, j, k, n, nq, ierr
  INTEGER(C_INT) :: cstat
  INTEGER(C_INT64_T) :: nw
  INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: wptr(:), wcell(:)
  REAL(C_DOUBLE), ALLOCATABLE, TARGET :: ww(:), wid(:), pc(:,:), pp(:,:)
  REAL(C_DOUBLE), ALLOCATABLE, TARGET :: wc(:), wx(:), wy(:), cvh(:), cvqq(:), cvin(:), cvout(:)
  INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: cvn(:), cvc1(:), cvc2(:)

  ! Options the GPU path implements (everything else stops the run).
  IF (opt_solver /= 2) CALL byebye('storm_fvt: only the FVT solver is provided.')
  IF (opt_friction /= 0) CALL byebye('storm_fvt: only Manning friction is provided.')
  IF (nctr2vtx /= 3) CALL byebye('storm_fvt: only inverse-distance vertex weights.')

  ! Module arrays the rest of the program reads (dataprepFVT.f90:31-62).
  ALLOCATE(z(n_elems), zvtx(n_pts), h(n_elems+1), hvtx(n_pts), u(n_elems), &
           uvtx(n_pts), v(n_elems), vvtx(n_pts), u_mag(n_elems), wavec(n_elems), &
           zeta(n_elems), zetavtx(n_pts), cdvtx(n_pts), cd(n_elems), rough(n_elems), &
           delh(n_elems), delu(n_elems), delv(n_elems), delz(n_elems), ev(n_pts), &
           phi(n_elems,3), e_est(n_elems), bdrycell(n_elems), drycell(n_elems), &
           partdry(n_elems), STAT=ierr)
  IF (ierr /= 0) CALL alloc_err(ierr)
  ALLOCATE(rkh(n_elems,0:rkorder), rku(n_elems,0:rkorder), rkv(n_elems,0:rkorder), STAT=ierr)
  IF (ierr /= 0) CALL alloc_err(ierr)
  uvtx = zero; vvtx = zero; cdvtx = zero; e_est = zero; ev = zero; phi = zero
  drycell = .FALSE.; bdrycell = .FALSE.; partdry = .FALSE.

  ! Mesh, 0-based (tecplot_read.f90 layout).
  ALLOCATE(xs(n_pts), ys(n_pts), tri(3*n_elems), ep(2*n_edges), ee(2*n_edges))
  DO i = 1, n_pts
    xs(i) = nodes(i)%x
    ys(i) = nodes(i)%y
  END DO
  DO i = 1, n_elems
    DO j = 1, 3
      tri(3*(i-1)+j) = grid(i)%vertex(j) - 1
    END DO
  END DO
  DO k = 1, n_edges
    ep(2*k-1) = edges(k)%p(1) - 1
    ep(2*k) = edges(k)%p(2) - 1
    ee(2*k-1) = MERGE(edges(k)%e(1) - 1, -1, edges(k)%e(1) >= 1)
    ee(2*k) = MERGE(edges(k)%e(2) - 1, -1, edges(k)%e(2) >= 1)
  END DO

  ! Initial data where tecplot_read left it (datav flags, dataprepFVT.f90:73-140).
  n = MERGE(n_pts, n_elems, datav(3))
  ALLOCATE(zs(n)); zs = z_tcp(1:n)
  n = MERGE(n_pts, n_elems, datav(4))
  ALLOCATE(hs(n)); hs = h_tcp(1:n)
  n = MERGE(n_pts, n_elems, datav(5))
  ALLOCATE(us(n)); us = u_tcp(1:n)
  n = MERGE(n_pts, n_elems, datav(6))
  ALLOCATE(vs(n)); vs = v_tcp(1:n)
  n = MERGE(n_pts, n_elems, datav(7))
  ALLOCATE(cds(n)); cds = cd_tcp(1:n)

  ! Boundary strings and their time tables (ioflow.f90).
  ALLOCATE(in_type(MAX(1,n_inflowbdr)), in_ptr(n_inflowbdr+1), qts_ptr(n_inflowbdr+1))
  in_ptr(1) = 0; qts_ptr(1) = 0
  DO k = 1, n_inflowbdr
    in_type(k) = vtype(k)
    in_ptr(k+1) = in_ptr(k) + n_qin(k)
    qts_ptr(k+1) = qts_ptr(k) + n_timeserq(k)
  END DO
  ALLOCATE(in_nodes(MAX(1,in_ptr(n_inflowbdr+1))), qts_t(MAX(1,qts_ptr(n_inflowbdr+1))), &
           qts_v(MAX(1,qts_ptr(n_inflowbdr+1))))
  DO k = 1, n_inflowbdr
    DO i = 1, n_qin(k)
      in_nodes(in_ptr(k)+i) = qin_nodes(i,k) - 1
    END DO
    DO i = 1, n_timeserq(k)
      qts_t(qts_ptr(k)+i) = timelocq(i,k)
      qts_v(qts_ptr(k)+i) = timeserq(i,k)
    END DO
  END DO
  ALLOCATE(out_type(MAX(1,n_outflowbdr)), out_ptr(n_outflowbdr+1), hts_ptr(n_outflowbdr+1))
  out_ptr(1) = 0; hts_ptr(1) = 0
  DO k = 1, n_outflowbdr
    out_type(k) = htype(k)
    out_ptr(k+1) = out_ptr(k) + n_hbc(k)
    nq = 0
    IF (htype(k) == 1) nq = n_timeserh(k)
    hts_ptr(k+1) = hts_ptr(k) + nq
  END DO
  ALLOCATE(out_nodes(MAX(1,out_ptr(n_outflowbdr+1))), hts_t(MAX(1,hts_ptr(n_outflowbdr+1))), &
           hts_v(MAX(1,hts_ptr(n_outflowbdr+1))))
  DO k = 1, n_outflowbdr
    DO i = 1, n_hbc(k)
      out_nodes(out_ptr(k)+i) = hbc_nodes(i,k) - 1
    END DO
    DO i = 1, hts_ptr(k+1) - hts_ptr(k)
      hts_t(hts_ptr(k)+i) = timeloch(i,k)
      hts_v(hts_ptr(k)+i) = timeserh(i,k)
    END DO
  END DO

  m%n_pts = n_pts; m%n_elems = n_elems; m%n_edges = n_edges
  m%x = C_LOC(xs); m%y = C_LOC(ys); m%tri = C_LOC(tri)
  m%edge_p = C_LOC(ep); m%edge_e = C_LOC(ee)
  m%z_at_vtx = MERGE(1, 0, datav(3)); m%h_at_vtx = MERGE(1, 0, datav(4))
  m%u_at_vtx = MERGE(1, 0, datav(5)); m%v_at_vtx = MERGE(1, 0, datav(6))
  m%cd_at_vtx = MERGE(1, 0, datav(7))
  m%z = C_LOC(zs); m%h = C_LOC(hs); m%u = C_LOC(us); m%v = C_LOC(vs); m%cd = C_LOC(cds)
  m%n_in = n_inflowbdr
  m%in_type = C_LOC(in_type); m%in_ptr = C_LOC(in_ptr); m%in_nodes = C_LOC(in_nodes)
  m%qts_ptr = C_LOC(qts_ptr); m%qts_t = C_LOC(qts_t); m%qts_v = C_LOC(qts_v)
  m%n_out = n_outflowbdr
  m%out_type = C_LOC(out_type); m%out_ptr = C_LOC(out_ptr); m%out_nodes = C_LOC(out_nodes)
  m%hts_ptr = C_LOC(hts_ptr); m%hts_t = C_LOC(hts_t); m%hts_v = C_LOC(hts_v)

  CALL storm_fvt_default_opts(o)
  o%rkorder = rkorder
  o%dt_op = MERGE(0, 1, dt_op == 0)
  o%steady = MERGE(1, 0, steady)
  o%stop_cr = stop_cr
  o%istop = istop
  o%btype = btype
  o%hvalve = MERGE(1, 0, hvalve)
  o%limiter = opt_limiter
  o%gradient = MERGE(2, 0, activateHDLS)  ! STORM_GRAD_HDLS or STORM_GRAD_LSQ
  o%vclip = MERGE(1, 0, vclip)
  o%cfl = cfl; o%rstop = rstop; o%h_dry = h_dry; o%h_wet = h_wet; o%vceiling = vceiling
  o%rsavg = MERGE(1, 0, activateRSAVG)
  o%rsavg_n = nRSAVG
  o%rsavg_eps = epsRSAVG

  cstat = storm_fvt_create(m, o, ctx)
  IF (cstat /= STORM_OK) CALL byebye('storm_fvt_create failed (mesh, options or device).')

  ! Wind data (dataprepFVT.f90:237-243; the file format of windforcing.f90:
  ! the count, then "c wx wy" per cell or per vertex).
  IF (wind_terms) THEN
    CALL read_wind(TRIM(wind_file), wc, wx, wy)
    cstat = storm_fvt_set_wind(ctx, INT(SIZE(wc), C_INT64_T), C_LOC(wc), C_LOC(wx), C_LOC(wy))
    IF (cstat /= STORM_OK) THEN
      PRINT *, TRIM(storm_fvt_message(ctx))
      CALL byebye('storm_fvt: invalid wind data.')
    END IF
  END IF

  ! Culverts (dataprepFVT.f90:309): ioculvert's data, located by the library.
  IF (culvert) THEN
    ALLOCATE(cvin(3*nculvert), cvout(2*nculvert), cvn(nculvert), cvc1(nculvert), &
             cvc2(nculvert), cvh(SUM(ncvtable(1:nculvert))), cvqq(SUM(ncvtable(1:nculvert))))
    k = 0
    DO i = 1, nculvert
      cvin(3*i-2:3*i) = cvptoin(1:3,i)
      cvout(2*i-1:2*i) = cvptoout(1:2,i)
      cvn(i) = ncvtable(i)
      DO j = 1, ncvtable(i)
        k = k + 1
        cvh(k) = cvhead(j,i)
        cvqq(k) = cvq(j,i)
      END DO
    END DO
    cstat = storm_fvt_set_culverts(ctx, INT(nculvert, C_INT32_T), C_LOC(cvin), C_LOC(cvout), &
                                   C_LOC(cvn), C_LOC(cvh), C_LOC(cvqq))
    IF (cstat == STORM_OK) cstat = storm_fvt_get_culverts(ctx, C_LOC(cvc1), C_LOC(cvc2), C_NULL_PTR)
    IF (cstat /= STORM_OK) THEN
      PRINT *, TRIM(storm_fvt_message(ctx))
      CALL byebye('storm_fvt: culvert set-up failed.')
    END IF
    cvtrigin(1:nculvert) = cvc1 + 1
    cvtrigout(1:nculvert) = cvc2 + 1
  END IF

  ! Module copies of the cell and vertex data (dataprepFVT.f90:69-140) and
  ! the centre-to-vertex weights of c2v_setup (type 3), as the library formed
  ! them (kept for later ctr2vtx calls of the output writers).
  nw = storm_fvt_get_c2v_weights(ctx, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR)
  IF (nw < 0) CALL byebye('storm_fvt: reading the set-up fields failed.')
  ALLOCATE(wptr(n_pts+1), wcell(MAX(1_C_INT64_T, nw)), ww(MAX(1_C_INT64_T, nw)), &
           wid(MAX(1_C_INT64_T, nw)), pc(n_elems,5), pp(n_pts,6), STAT=ierr)
  IF (ierr /= 0) CALL alloc_err(ierr)
  nw = storm_fvt_get_c2v_weights(ctx, C_LOC(wptr), C_LOC(wcell), C_LOC(ww), C_LOC(wid))
  cstat = storm_fvt_get_prep_fields(ctx, C_LOC(pc(1,1)), C_LOC(pp(1,1)), C_LOC(pc(1,2)), &
                                    C_LOC(pc(1,3)), C_LOC(pp(1,2)))
  IF (cstat == STORM_OK) cstat = storm_fvt_get_state(ctx, C_LOC(pc(1,4)), C_LOC(pc(1,5)), &
                                                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR)
  IF (cstat == STORM_OK) cstat = storm_fvt_get_vertex(ctx, C_LOC(pp(1,3)), C_LOC(pp(1,4)), &
                                                      C_LOC(pp(1,5)), C_LOC(pp(1,6)))
  IF (nw < 0 .OR. cstat /= STORM_OK) CALL byebye('storm_fvt: reading the set-up fields failed.')
  z = pc(:,1)
  zeta = pc(:,2)
  cd = pc(:,3)
  h(1:n_elems) = pc(:,4)
  h(n_elems+1) = zero
  u = pc(:,5)
  zvtx = pp(:,1)
  cdvtx = pp(:,2)
  hvtx = pp(:,3)
  uvtx = pp(:,4)
  vvtx = pp(:,5)
  zetavtx = pp(:,6)
  cstat = storm_fvt_get_state(ctx, C_NULL_PTR, C_NULL_PTR, C_LOC(pc(1,5)), C_NULL_PTR, C_NULL_PTR)
  IF (cstat /= STORM_OK) CALL byebye('storm_fvt: reading the set-up fields failed.')
  v = pc(:,5)
  k = MAX(2, SIZE(n2t2, 2))
  ALLOCATE(c2v_weights(n_pts, k-1), c2v_weights_id(n_pts, k-1), STAT=ierr)
  IF (ierr /= 0) CALL alloc_err(ierr)
  c2v_weights = zero
  c2v_weights_id = zero
  DO i = 1, n_pts
    DO j = 1, wptr(i+1) - wptr(i)
      c2v_weights(i, j) = ww(wptr(i) + j)
      c2v_weights_id(i, j) = wid(wptr(i) + j)
    END DO
  END DO
  rough = cd
  DEALLOCATE(z_tcp, h_tcp, u_tcp, v_tcp, cd_tcp)
  datav = .FALSE.
  datav(1) = .TRUE.; datav(2) = .TRUE.; datav(3) = .TRUE.

CONTAINS

  ! The wind file: its count, then one "c wx wy" line per value; blank lines
  ! and lines starting with '#' are skipped (readline.f90's rules).
  SUBROUTINE read_wind(fname, c, a, b)
    CHARACTER(LEN=*), INTENT(IN) :: fname
    REAL(C_DOUBLE), ALLOCATABLE, INTENT(OUT) :: c(:), a(:), b(:)
    CHARACTER(LEN=256) :: line
    INTEGER :: u, ios, nv, iv
    OPEN (NEWUNIT=u, FILE=fname, STATUS='OLD', ACTION='READ', IOSTAT=ios)
    IF (ios /= 0) CALL byebye('storm_fvt: unable to open the wind file.')
    CALL next_line(u, line)
    READ (line, *, IOSTAT=ios) nv
    IF (ios /= 0 .OR. nv < 1) CALL byebye('storm_fvt: invalid wind file.')
    ALLOCATE(c(nv), a(nv), b(nv))
    DO iv = 1, nv
      CALL next_line(u, line)
      READ (line, *, IOSTAT=ios) c(iv), a(iv), b(iv)
      IF (ios /= 0) CALL byebye('storm_fvt: invalid wind file.')
    END DO
    CLOSE (u)
  END SUBROUTINE read_wind

  SUBROUTINE next_line(u, line)
    INTEGER, INTENT(IN) :: u
    CHARACTER(LEN=*), INTENT(OUT) :: line
    INTEGER :: ios
    DO
      READ (u, '(A)', IOSTAT=ios) line
      IF (ios /= 0) CALL byebye('storm_fvt: the wind file ends early.')
      line = ADJUSTL(line)
      IF (line(1:1) == '#' .OR. LEN_TRIM(line) == 0) CYCLE
      RETURN
    END DO
  END SUBROUTINE next_line
END SUBROUTINE dataprepFVT


SUBROUTINE solverFVT(iter, ttime, rmax, fn, title)
  USE, INTRINSIC :: ISO_C_BINDING
  USE parameters
  USE geometry
  USE dep_vars
  USE options
  USE RKparams
  USE constants
  USE vbc_arrays
  USE storm_fvt_c
  USE storm_fvt_dropin_state
  IMPLICIT NONE
  INTEGER, INTENT(OUT) :: iter
  REAL(KIND=mp), INTENT(OUT) :: rmax, ttime
  CHARACTER (LEN=*), INTENT(IN) :: fn, title

  INTEGER(C_INT32_T) :: it
  REAL(C_DOUBLE) :: tt, rm
  REAL(C_DOUBLE), ALLOCATABLE, TARGET :: hh(:), uu(:), vv(:), ph(:), qh(:), qu(:), qv(:), &
                                         hv(:), zv(:), uv(:), vv2(:), cv(:), ee(:), gx(:), &
                                         gy(:), qb(:), hb(:), cs(:)
  INTEGER(C_INT8_T), ALLOCATABLE, TARGET :: dr(:)
  INTEGER(C_INT) :: cstat
  INTEGER :: i, k

  IF (.NOT. C_ASSOCIATED(ctx)) CALL byebye('solverFVT: dataprepFVT was not called.')
  cstat = storm_fvt_run(ctx, INT(istop, C_INT32_T), it, tt, rm)
  IF (cstat == STORM_OK) cstat = storm_fvt_finalize(ctx)
  IF (cstat /= STORM_OK) THEN
    PRINT *, TRIM(storm_fvt_message(ctx))
    CALL byebye('solverFVT: the GPU time loop failed.')
  END IF
  iter = it
  ttime = tt
  rmax = rm

  ALLOCATE(hh(n_elems), uu(n_elems), vv(n_elems), ph(3*n_elems), dr(n_elems), &
           qh(n_elems), qu(n_elems), qv(n_elems), hv(n_pts), zv(n_pts), uv(n_pts), &
           vv2(n_pts), cv(n_pts), ee(n_elems), gx(n_elems), gy(n_elems), &
           qb(MAX(1, n_inflowbdr)), hb(MAX(1, n_outflowbdr)))
  cstat = storm_fvt_get_state(ctx, C_LOC(hh), C_LOC(uu), C_LOC(vv), C_LOC(ph), C_LOC(dr))
  IF (cstat == STORM_OK) cstat = storm_fvt_get_conserved(ctx, C_LOC(qh), C_LOC(qu), C_LOC(qv))
  IF (cstat == STORM_OK) cstat = storm_fvt_get_vertex(ctx, C_LOC(hv), C_LOC(uv), C_LOC(vv2), &
                                                      C_LOC(zv))
  IF (cstat == STORM_OK) cstat = storm_fvt_get_output(ctx, C_LOC(cv), C_LOC(ee))
  IF (cstat == STORM_OK) cstat = storm_fvt_get_gradients(ctx, C_LOC(gx), C_LOC(gy))
  IF (cstat == STORM_OK) cstat = storm_fvt_get_bc_values(ctx, C_LOC(qb), C_LOC(hb))
  IF (cstat == STORM_OK .AND. culvert) THEN  ! cvdisch's discharges of the last stage
    ALLOCATE(cs(nculvert))
    cstat = storm_fvt_get_culverts(ctx, C_NULL_PTR, C_NULL_PTR, C_LOC(cs))
    IF (cstat == STORM_OK) cvsrc(1:nculvert) = cs
  END IF
  IF (cstat /= STORM_OK) CALL byebye('solverFVT: reading the GPU state failed.')
  h(1:n_elems) = hh
  u = uu
  v = vv
  DO i = 1, n_elems
    phi(i,1) = ph(i)
    phi(i,2) = ph(n_elems + i)
    phi(i,3) = ph(2*n_elems + i)
    drycell(i) = dr(i) /= 0
    delh(i)%x = gx(i)
    delh(i)%y = gy(i)
  END DO
  rkh(:,rkorder) = qh
  rku(:,rkorder) = qu
  rkv(:,rkorder) = qv
  hvtx = hv
  uvtx = uv
  vvtx = vv2
  zetavtx = zv
  cdvtx = cv
  e_est = ee
  zeta = z + h(1:n_elems)
  ! the boundary values of the last step (get_time_bcs, solverFVT.f90:462)
  DO k = 1, n_inflowbdr
    qin(k) = qb(k)
  END DO
  DO k = 1, n_outflowbdr
    IF (htype(k) == 1) hbc(k) = hb(k)
  END DO
  tecplot_print = .TRUE.
  IF (LEN(fn) < 0 .OR. LEN(title) < 0) RETURN  ! names kept for the call surface
END SUBROUTINE solverFVT
