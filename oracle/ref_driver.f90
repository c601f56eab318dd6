! Test infrastructure only: a driver that runs the reference SToRM FVT path
! (compiled from the reference's own Fortran sources by oracle/Makefile) on a
! ".run" case and dumps the final cell state in raw little-endian binary.
!
! It follows the reference's own main-program sequence for a ".run" case,
! Source_xsy-64bit/SToRM.f90:64-66 -> read_data.f90:80-83 (read_mfa) ->
! read_data.f90:92 (ioculvert, when the options name a culvert file) ->
! solver.f90:50,57 (dataprepFVT, solverFVT), without the command-line prompt
! of read_data.f90:41-52 and without the final Tecplot/CGNS writers.
!
! Usage: ref_driver <case.run> <out.bin> [prep.bin]
!
! With a third argument the module copies dataprepFVT leaves are written to
! it right after dataprepFVT (same encoding):
!   real64 z(N), zvtx(P), zeta(N), cd(N), cdvtx(P), h(N), u(N), v(N),
!          hvtx(P), uvtx(P), vvtx(P), zetavtx(P)
!
! The wall time of the solverFVT call alone is printed on standard output as
! "solverFVT seconds <t>" (the CPU baseline of the benchmark reads it).
!
! Output layout (stream access, native endianness):
!   int32  n_elems, n_pts, iter, rkorder
!   real64 ttime, rmax
!   real64 h(N), u(N), v(N)                 after the final to_vertices call
!   real64 rkh(N), rku(N), rkv(N)           the RK slot rkorder (q^{n+1})
!   real64 phi(N,1), phi(N,2), phi(N,3)
!   int32  drycell(N)                       1 = dry
!   real64 hvtx(P), zetavtx(P)
!   real64 uvtx(P), vvtx(P), cdvtx(P), zvtx(P)
!   real64 delh(N)%x, delh(N)%y, e_est(N)     what error_est(2) read and wrote
PROGRAM ref_driver
  USE parameters
  USE geometry
  USE dep_vars
  USE options
  USE io
  USE RKparams
  IMPLICIT NONE
  CHARACTER (LEN=250) :: fname, oname, pname
  CHARACTER (LEN=160) :: title
  INTEGER :: iter, i, u_out
  INTEGER (KIND=8) :: c0, c1, crate
  INTEGER, ALLOCATABLE :: dryi(:)
  REAL (KIND=mp) :: ttime, rmax

  CALL GET_COMMAND_ARGUMENT(1, fname)
  CALL GET_COMMAND_ARGUMENT(2, oname)
  cgns = .FALSE.
  CALL read_mfa(fname, title)
  IF (culvert) CALL ioculvert   ! read_data.f90:92
  CALL dataprepFVT
  IF (COMMAND_ARGUMENT_COUNT() >= 3) THEN
    CALL GET_COMMAND_ARGUMENT(3, pname)
    OPEN (NEWUNIT=u_out, FILE=pname, ACCESS='STREAM', FORM='UNFORMATTED', &
          STATUS='REPLACE')
    WRITE (u_out) z(1:n_elems), zvtx(1:n_pts), zeta(1:n_elems), cd(1:n_elems), &
                  cdvtx(1:n_pts), h(1:n_elems), u(1:n_elems), v(1:n_elems), &
                  hvtx(1:n_pts), uvtx(1:n_pts), vvtx(1:n_pts), zetavtx(1:n_pts)
    CLOSE (u_out)
  END IF
  CALL SYSTEM_CLOCK(c0, crate)
  CALL solverFVT(iter, ttime, rmax, fname, title)
  CALL SYSTEM_CLOCK(c1)
  WRITE (*, '(A,ES16.8)') 'solverFVT seconds ', REAL(c1 - c0, mp) / REAL(crate, mp)

  ALLOCATE(dryi(n_elems))
  DO i = 1, n_elems
    dryi(i) = 0
    IF (drycell(i)) dryi(i) = 1
  END DO
  OPEN (NEWUNIT=u_out, FILE=oname, ACCESS='STREAM', FORM='UNFORMATTED', &
        STATUS='REPLACE')
  WRITE (u_out) n_elems, n_pts, iter, rkorder
  WRITE (u_out) ttime, rmax
  WRITE (u_out) h(1:n_elems), u(1:n_elems), v(1:n_elems)
  WRITE (u_out) rkh(1:n_elems, rkorder), rku(1:n_elems, rkorder), &
                rkv(1:n_elems, rkorder)
  WRITE (u_out) phi(1:n_elems, 1), phi(1:n_elems, 2), phi(1:n_elems, 3)
  WRITE (u_out) dryi(1:n_elems)
  WRITE (u_out) hvtx(1:n_pts), zetavtx(1:n_pts)
  WRITE (u_out) uvtx(1:n_pts), vvtx(1:n_pts), cdvtx(1:n_pts), zvtx(1:n_pts)
  WRITE (u_out) delh(1:n_elems)%x, delh(1:n_elems)%y, e_est(1:n_elems)
  CLOSE (u_out)
END PROGRAM ref_driver
