! Test infrastructure only: calls single routines of the reference SToRM
! (compiled from the reference's own sources by oracle/Makefile) on inputs
! read from a file and writes their outputs, for the per-routine parity tests
! of the oracle and of the library.  Built and run in this container only
! (tests/golden/make_units.py), never shipped.
!
! Usage: ref_units <mode> <in.bin> <out.bin> [case.run]
!   rusanov  in:  int32 K; real64 r(10,K) = (nx,ny,hl,ul,vl,cl,hr,ur,vr,cr)
!            out: real64 f(3,K)                 flux_rusanov.f90:1-82
!   hll      in:  int32 K; real64 r(9,K) = (nx,ny,hl,ul,vl,hr,ur,vr,h0)
!            out: real64 f(3,K)                 flux_HLL.f90:1-112 (the caller
!            sends only states whose waves both run one way: the star
!            region of this routine is defective)
!   fric     in:  int32 K; int32 dry(K); real64 cd, h, umag, u, v, area,
!                 phi2, phi3 (K each)
!            out: real64 phi2(K), phi3(K)       fric_terms.f90:22-98, Manning
!   rds      the mesh of case.run (read_mfa); in: real64 h_dry, z(P), h(P)
!            out: real64 source(P,3)            fvsources.f90:27-46 with the
!                 node least squares of lsq_setup(1,2) (dataprepRDS.f90:93)
!                 real64 zout(3,N)              drynwet.f90:28-50 per triangle,
!                 int32 pd(4,N)                 part_dry.f90:30-43: the flag,
!                                               n, i, j (i, j 0 where unset)
PROGRAM ref_units
  USE parameters
  USE geometry
  USE dep_vars
  USE options
  USE io
  IMPLICIT NONE
  CHARACTER(LEN=250) :: mode, iname, oname, rname
  CHARACTER(LEN=160) :: title
  INTEGER :: k, i, nk, u_in, u_out, n, pi, pj
  INTEGER, ALLOCATABLE :: di(:), pd(:,:)
  REAL(KIND=mp), ALLOCATABLE :: r(:,:), f(:,:), a(:,:), zo(:,:)
  REAL(KIND=mp) :: fl(3), zz(3), hd
  TYPE(edge) :: e
  LOGICAL, EXTERNAL :: part_dry

  CALL GET_COMMAND_ARGUMENT(1, mode)
  CALL GET_COMMAND_ARGUMENT(2, iname)
  CALL GET_COMMAND_ARGUMENT(3, oname)
  OPEN (NEWUNIT=u_in, FILE=iname, ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  OPEN (NEWUNIT=u_out, FILE=oname, ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  e%p = 0
  e%e = 0
  e%tang = 0.0_mp
  e%length = 1.0_mp

  SELECT CASE (TRIM(mode))
  CASE ('rusanov')
    READ (u_in) nk
    ALLOCATE(r(10,nk), f(3,nk))
    READ (u_in) r
    DO k = 1, nk
      e%normal(1) = r(1,k)
      e%normal(2) = r(2,k)
      CALL flux_rusanov(e, r(3,k), r(4,k), r(5,k), r(6,k), r(7,k), r(8,k), r(9,k), &
                        r(10,k), fl)
      f(:,k) = fl
    END DO
    WRITE (u_out) f

  CASE ('hll')
    READ (u_in) nk
    ALLOCATE(r(9,nk), f(3,nk))
    READ (u_in) r
    DO k = 1, nk
      e%normal(1) = r(1,k)
      e%normal(2) = r(2,k)
      CALL flux_HLL(e, r(3,k), r(4,k), r(5,k), r(6,k), r(7,k), r(8,k), r(9,k), fl)
      f(:,k) = fl
    END DO
    WRITE (u_out) f

  CASE ('fric')
    READ (u_in) nk
    n_elems = nk
    ALLOCATE(di(nk), a(nk,8), drycell(nk), cd(nk), h(nk), u_mag(nk), u(nk), v(nk), &
             rough(nk), phi(nk,3), grid(nk))
    READ (u_in) di
    READ (u_in) a
    opt_friction = 0
    DO k = 1, nk
      drycell(k) = di(k) /= 0
      cd(k) = a(k,1); h(k) = a(k,2); u_mag(k) = a(k,3); u(k) = a(k,4); v(k) = a(k,5)
      grid(k)%area = a(k,6)
      phi(k,1) = 0.0_mp; phi(k,2) = a(k,7); phi(k,3) = a(k,8)
    END DO
    CALL fric_terms
    WRITE (u_out) phi(:,2), phi(:,3)

  CASE ('rds')
    CALL GET_COMMAND_ARGUMENT(4, rname)
    cgns = .FALSE.
    CALL read_mfa(rname, title)
    READ (u_in) hd
    ALLOCATE(z(n_pts), h(n_pts), zbx(n_pts), zby(n_pts), source(n_pts,3), &
             zo(3,n_elems), pd(4,n_elems))
    READ (u_in) z
    READ (u_in) h
    h_dry = hd
    zadjust = .TRUE.
    CALL lsq_setup(1, 2)
    CALL fvsources
    DO k = 1, n_elems
      CALL drynwet(grid(k), zz)
      zo(:,k) = zz
      n = 0; pi = 0; pj = 0
      pd(1,k) = MERGE(1, 0, part_dry(grid(k), n, pi, pj))
      pd(2,k) = n
      pd(3,k) = MERGE(pi, 0, n >= 1)
      pd(4,k) = MERGE(pj, 0, n >= 1)
    END DO
    WRITE (u_out) source(:,1), source(:,2), source(:,3)
    WRITE (u_out) zo
    WRITE (u_out) pd

  CASE DEFAULT
    PRINT *, 'ref_units: unknown mode ', TRIM(mode)
    STOP 1
  END SELECT
  CLOSE (u_in)
  CLOSE (u_out)
  i = 0
END PROGRAM ref_units
