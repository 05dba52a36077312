!==============================================================================================
! ref_driver.f90 -- TEST INFRASTRUCTURE ONLY.  Never linked into, loaded by or called from the
! product path.  It is linked against the reference Fortran modules compiled in place from
! /root/reference (see oracle/Makefile) and exposes their public routines to the test suite
! through raw stream files (little-endian, no record markers):
!
!   ref_driver lambda <in> <out>
!       compute_lambda_vacc (SOURCES/condlim.f90:879-957) on a list of (U_l, U_r, n) triples.
!   ref_driver fct <in> <out>
!       FCT_positivity (SOURCES/fct.f90:5-42) on a given CSR matrix of antidiffusive fluxes.
!   ref_driver setup <data> <init|none> <outdir>
!       read_my_data + construct_mesh + construct_matrices + init (SOURCES/main.f90:19-40) and
!       dump the mesh, CSR pattern, mass, lumped mass, diag index, c_ij, bathymetry, U0 and the
!       Dirichlet node lists.
!   ref_driver dij <data> <init|none> <state> <out>
!       compute_dij (SOURCES/update.f90:291-326) at a given state; writes dij%aa.
!   ref_driver euler <data> <init|none> <state> <dt> <outdir>
!       one forward-Euler stage, euler(un,unext) (SOURCES/update.f90:134-289); writes unext
!       and dij%aa as left by euler.
!   ref_driver run <data> <init|none> <nsteps> <outdir> [dump_every]
!       (states are un(syst_size,np): 3 components, 5 for the SGN model of test 13)
!       nsteps SSP-RK3 steps as in the time loop of SOURCES/main.f90:61-92 (fixed step count
!       instead of the Tfinal loop); writes dt per step, U after each stage of step 1, the final
!       U, optional snapshots and the wall time of the loop.
!   ref_driver errors <data> <init|none> <state> <time> <out>
!       the quadrature norms of compute_errors (SOURCES/main.f90:198-263) at a given state and
!       time: for each component k = 1..3 of sol_anal (SOURCES/condlim.f90:311-831) at the
!       Gauss points, ns_anal_l1 and ns_anal_0 (SFEMaNS/FEMSUB_OBJET/fem_tn.f90:1513-1553,
!       1679-1716) of U(k,:) - exact and of 0 - exact; writes 12 reals (k-major: L1 error, L1 norm,
!       L2 error, L2 norm).
!   ref_driver flow <data> <init|none> <state> <out>
!       the diagnostic the time loop writes to unit 10 for tests 1, 2 and 10 (SOURCES/main.f90:94-99):
!       ns_l1 (SFEMaNS/FEMSUB_OBJET/fem_tn.f90:2119-2153) of U(2,:), U(3,:) and U(1,:)*SQRT(U(1,:)),
!       and 0.5*(q1+q2)/(gravity*q3); writes those 4 reals.
!
! <init> is an optional custom initial-data file for problems that have no init case in the
! reference (see custom_init below); 'none' calls the reference init (SOURCES/condlim.f90:92).
!==============================================================================================
PROGRAM ref_driver
  USE space_dim
  USE input_data
  USE mesh_handling
  USE update
  USE boundary_conditions
  USE fct
  USE matrix_type
  USE fem_tn
  USE iso_c_binding, ONLY : c_double
  IMPLICIT NONE
  INTERFACE
     ! sin rounded to nearest, from the C checker (swe_oracle.c): the inflow value's definition
     FUNCTION orc_sin_rn(x) BIND(C, NAME='orc_sin_rn') RESULT(s)
       IMPORT :: c_double
       REAL(c_double), VALUE :: x
       REAL(c_double) :: s
     END FUNCTION orc_sin_rn
  END INTERFACE
  CHARACTER(LEN=512) :: cmd
  INTEGER :: bc_kind = 0
  REAL(KIND=8) :: bc_amp = 0.d0, bc_period = 1.d0, bc_level = 0.d0
  REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE :: un

  CALL GET_COMMAND_ARGUMENT(1, cmd)
  SELECT CASE (TRIM(cmd))
  CASE ('lambda')
     CALL do_lambda
  CASE ('fct')
     CALL do_fct
  CASE ('setup')
     CALL do_setup
  CASE ('dij')
     CALL do_dij
  CASE ('euler')
     CALL do_euler
  CASE ('run')
     CALL do_run
  CASE ('errors')
     CALL do_errors
  CASE ('flow')
     CALL do_flow
  CASE DEFAULT
     WRITE(*,*) 'ref_driver: unknown command ', TRIM(cmd)
     STOP 2
  END SELECT

CONTAINS

  FUNCTION arg(k) RESULT(s)
    INTEGER, INTENT(IN) :: k
    CHARACTER(LEN=512)  :: s
    CALL GET_COMMAND_ARGUMENT(k, s)
  END FUNCTION arg

  SUBROUTINE open_in(u, path)
    INTEGER, INTENT(IN) :: u
    CHARACTER(LEN=*), INTENT(IN) :: path
    OPEN(UNIT=u, FILE=TRIM(path), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
  END SUBROUTINE open_in

  SUBROUTINE open_out(u, path)
    INTEGER, INTENT(IN) :: u
    CHARACTER(LEN=*), INTENT(IN) :: path
    OPEN(UNIT=u, FILE=TRIM(path), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE', ACTION='WRITE')
  END SUBROUTINE open_out

  !---------------------------------------------------------------------------------------------
  SUBROUTINE do_lambda
    INTEGER :: n, tt, i
    REAL(KIND=8) :: g, ht, lambda
    REAL(KIND=8), DIMENSION(8) :: rec
    REAL(KIND=8), DIMENSION(3) :: ul, ur
    REAL(KIND=8), DIMENSION(2) :: nij
    CALL open_in(40, arg(2))
    READ(40) n, tt, g, ht
    inputs%syst_size = 3
    inputs%type_test = tt
    inputs%gravity = g
    inputs%htiny = ht
    CALL open_out(41, arg(3))
    DO i = 1, n
       READ(40) rec
       ul = rec(1:3)
       ur = rec(4:6)
       nij = rec(7:8)
       CALL compute_lambda_vacc(ul, ur, nij, lambda)
       WRITE(41) lambda
    END DO
    CLOSE(40)
    CLOSE(41)
  END SUBROUTINE do_lambda

  !---------------------------------------------------------------------------------------------
  SUBROUTINE do_fct
    INTEGER :: n, nnz
    REAL(KIND=8) :: ht
    TYPE(matrice_bloc) :: mat, lij
    REAL(KIND=8), DIMENSION(:), ALLOCATABLE :: unext
    CALL open_in(40, arg(2))
    READ(40) n, nnz, ht
    inputs%htiny = ht
    ALLOCATE(mat%ia(n+1), mat%ja(nnz), mat%aa(nnz), unext(n))
    ALLOCATE(lij%ia(n+1), lij%ja(nnz), lij%aa(nnz))
    READ(40) mat%ia, mat%ja, mat%aa, unext
    CLOSE(40)
    lij%ia = mat%ia
    lij%ja = mat%ja
    lij%aa = 0.d0
    CALL FCT_positivity(unext, mat, lij)
    CALL open_out(41, arg(3))
    WRITE(41) lij%aa
    CLOSE(41)
  END SUBROUTINE do_fct

  !---------------------------------------------------------------------------------------------
  ! Problem set-up as in SOURCES/main.f90:19-40.
  SUBROUTINE prepare(datafile, initfile)
    CHARACTER(LEN=*), INTENT(IN) :: datafile, initfile
    CALL read_my_data(TRIM(datafile))
    CALL construct_mesh
    CALL construct_matrices
    inputs%syst_size = k_dim+1
    IF (inputs%type_test==13) inputs%syst_size = k_dim+3   ! SGN model: h, hu, hv, h*eta, h*w
    ALLOCATE(un(inputs%syst_size, mesh%np))
    inputs%time = 0.d0
    IF (TRIM(initfile) == 'none') THEN
       CALL init(un)
    ELSE
       CALL custom_init(initfile)
    END IF
  END SUBROUTINE prepare

  ! Initial data for problems with no init case in the reference.  File layout:
  !   int32 np, int32 bc_kind, real8 gravity, max_water_h, bc_amp, bc_period, bc_level,
  !   real8 bath(np), real8 un(3,np)
  ! It allocates the same module arrays as init (SOURCES/condlim.f90:105-112) and sets htiny
  ! as init does (SOURCES/condlim.f90:233).
  SUBROUTINE custom_init(initfile)
    CHARACTER(LEN=*), INTENT(IN) :: initfile
    INTEGER :: np
    REAL(KIND=8) :: g, hmax
    CALL open_in(42, initfile)
    READ(42) np, bc_kind, g, hmax, bc_amp, bc_period, bc_level
    IF (np /= mesh%np) THEN
       WRITE(*,*) 'ref_driver: init file has np=', np, ' mesh has ', mesh%np
       STOP 3
    END IF
    ALLOCATE(bath(mesh%np), one_over_h(mesh%np), regul_h(mesh%np), velocity(k_dim, mesh%np))
    READ(42) bath, un
    CLOSE(42)
    inputs%gravity = g
    max_water_h = hmax
    inputs%htiny = inputs%epsilon*max_water_h
  END SUBROUTINE custom_init

  ! Dirichlet overwrite.  bc_kind 0 is the reference bdy (SOURCES/main.f90:176-188); kinds 1 and
  ! 2 give values for the problems without a sol_anal case: walls (q.n = 0 on the listed nodes)
  ! and walls plus a sinusoidal free-surface inflow on the h list (sin rounded to nearest).
  SUBROUTINE bdy_ref(uu, t)
    REAL(KIND=8), DIMENSION(:,:), INTENT(INOUT) :: uu
    REAL(KIND=8), INTENT(IN) :: t
    INTEGER :: i, n
    REAL(KIND=8) :: pi
    IF (bc_kind == 0) THEN
       IF (SIZE(h_js_D).NE.0)  uu(1,h_js_D)  = sol_anal(1,mesh%rr(:,h_js_D),t)
       IF (SIZE(ux_js_D).NE.0) uu(2,ux_js_D) = sol_anal(2,mesh%rr(:,ux_js_D),t)
       IF (SIZE(uy_js_D).NE.0) uu(3,uy_js_D) = sol_anal(3,mesh%rr(:,uy_js_D),t)
       IF (inputs%type_test==13) THEN
          IF (SIZE(ux_js_D).NE.0) uu(4,ux_js_D) = sol_anal(4,mesh%rr(:,ux_js_D),t)
          IF (SIZE(uy_js_D).NE.0) uu(5,uy_js_D) = sol_anal(5,mesh%rr(:,uy_js_D),t)
       END IF
       RETURN
    END IF
    pi = ACOS(-1.d0)
    IF (bc_kind == 2) THEN
       DO i = 1, SIZE(h_js_D)
          n = h_js_D(i)
          uu(1,n) = MAX(bc_level - bath(n) + bc_amp*orc_sin_rn(2.d0*pi*t/bc_period), 0.d0)
       END DO
    END IF
    DO i = 1, SIZE(ux_js_D)
       uu(2,ux_js_D(i)) = 0.d0
    END DO
    DO i = 1, SIZE(uy_js_D)
       uu(3,uy_js_D(i)) = 0.d0
    END DO
  END SUBROUTINE bdy_ref

  ! COMPUTE_DT of SOURCES/main.f90:166-174.
  SUBROUTINE compute_dt_ref(u0)
    REAL(KIND=8), DIMENSION(:,:), INTENT(IN) :: u0
    CALL compute_dij(u0)
    inputs%dt = inputs%CFL*1/MAXVAL(ABS(dij%aa(diag))/lumped)
  END SUBROUTINE compute_dt_ref

  SUBROUTINE read_state(path, uu)
    CHARACTER(LEN=*), INTENT(IN) :: path
    REAL(KIND=8), DIMENSION(:,:), INTENT(OUT) :: uu
    CALL open_in(43, path)
    READ(43) uu
    CLOSE(43)
  END SUBROUTINE read_state

  SUBROUTINE write_real(path, a)
    CHARACTER(LEN=*), INTENT(IN) :: path
    REAL(KIND=8), DIMENSION(:), INTENT(IN) :: a
    CALL open_out(44, path)
    WRITE(44) a
    CLOSE(44)
  END SUBROUTINE write_real

  SUBROUTINE write_state(path, a)
    CHARACTER(LEN=*), INTENT(IN) :: path
    REAL(KIND=8), DIMENSION(:,:), INTENT(IN) :: a
    CALL open_out(44, path)
    WRITE(44) a
    CLOSE(44)
  END SUBROUTINE write_state

  SUBROUTINE write_int(path, a)
    CHARACTER(LEN=*), INTENT(IN) :: path
    INTEGER, DIMENSION(:), INTENT(IN) :: a
    CALL open_out(44, path)
    WRITE(44) a
    CLOSE(44)
  END SUBROUTINE write_int

  !---------------------------------------------------------------------------------------------
  SUBROUTINE do_setup
    CHARACTER(LEN=512) :: od
    INTEGER :: m, ni, nj, i, j, d, p
    REAL(KIND=8), DIMENSION(k_dim) :: x
    REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE :: cij
    REAL(KIND=8), DIMENSION(6) :: scal
    CALL prepare(arg(2), arg(3))
    od = arg(4)
    ! c_ij is private to MODULE update; it is rebuilt here with the same element loop and the
    ! same SUM expression as SOURCES/update.f90:106-124 so the checker can read its bits.
    ALLOCATE(cij(k_dim, SIZE(mass%ja)))
    cij = 0.d0
    DO m = 1, mesh%me
       DO ni = 1, mesh%gauss%n_w
          i = mesh%jj(ni, m)
          DO nj = 1, mesh%gauss%n_w
             j = mesh%jj(nj, m)
             DO d = 1, k_dim
                x(d) = SUM(mesh%gauss%dw(d,nj,:,m) * mesh%gauss%ww(ni,:)*mesh%gauss%rj(:,m))
             END DO
             DO p = mass%ia(i), mass%ia(i+1) - 1
                IF (mass%ja(p) == j) THEN
                   cij(:,p) = cij(:,p) + x
                   EXIT
                END IF
             END DO
          END DO
       END DO
    END DO
    CALL write_int(TRIM(od)//'/dims.i4', [mesh%np, mesh%me, SIZE(mass%ja), SIZE(h_js_D), &
         SIZE(ux_js_D), SIZE(uy_js_D), inputs%type_test])
    scal = [inputs%gravity, inputs%htiny, max_water_h, inputs%CFL, inputs%epsilon, inputs%time]
    CALL write_real(TRIM(od)//'/scalars.f8', scal)
    CALL write_int(TRIM(od)//'/jj.i4', RESHAPE(mesh%jj, [SIZE(mesh%jj)]))
    CALL write_real(TRIM(od)//'/rr.f8', RESHAPE(mesh%rr, [SIZE(mesh%rr)]))
    CALL write_int(TRIM(od)//'/ia.i4', mass%ia)
    CALL write_int(TRIM(od)//'/ja.i4', mass%ja)
    CALL write_real(TRIM(od)//'/mass.f8', mass%aa)
    CALL write_real(TRIM(od)//'/lumped.f8', lumped)
    CALL write_int(TRIM(od)//'/diag.i4', diag)
    CALL write_real(TRIM(od)//'/cij.f8', RESHAPE(cij, [SIZE(cij)]))
    CALL write_real(TRIM(od)//'/bath.f8', bath)
    CALL write_state(TRIM(od)//'/u0.f8', un)
    CALL write_int(TRIM(od)//'/h_js_D.i4', h_js_D)
    CALL write_int(TRIM(od)//'/ux_js_D.i4', ux_js_D)
    CALL write_int(TRIM(od)//'/uy_js_D.i4', uy_js_D)
  END SUBROUTINE do_setup

  !---------------------------------------------------------------------------------------------
  SUBROUTINE do_dij
    CALL prepare(arg(2), arg(3))
    CALL read_state(arg(4), un)
    CALL compute_dij(un)
    CALL write_real(arg(5), dij%aa)
  END SUBROUTINE do_dij

  !---------------------------------------------------------------------------------------------
  SUBROUTINE do_euler
    REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE :: unext
    CHARACTER(LEN=512) :: od, s
    REAL(KIND=8) :: dt
    CALL prepare(arg(2), arg(3))
    CALL read_state(arg(4), un)
    s = arg(5)
    READ(s, *) dt
    od = arg(6)
    inputs%dt = dt
    ALLOCATE(unext(inputs%syst_size, mesh%np))
    CALL euler(un, unext)
    CALL write_state(TRIM(od)//'/unext.f8', unext)
    CALL write_real(TRIM(od)//'/dij.f8', dij%aa)
  END SUBROUTINE do_euler

  !---------------------------------------------------------------------------------------------
  ! The sums of compute_errors (SOURCES/main.f90:198-263).  The Gauss point coordinates are
  ! r_gauss's (SOURCES/main.f90:303-322, an internal procedure of the main program): the same
  ! element / Gauss point / coordinate loops and the same sum, written out here.
  SUBROUTINE do_errors
    CHARACTER(LEN=512) :: s
    REAL(KIND=8) :: t, r
    REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE :: rg
    REAL(KIND=8), DIMENSION(:), ALLOCATABLE :: uex, zero
    REAL(KIND=8), DIMENSION(12) :: out
    INTEGER :: idx, k, l, m, n
    CALL prepare(arg(2), arg(3))
    CALL read_state(arg(4), un)
    s = arg(5)
    READ(s, *) t
    inputs%time = t
    ALLOCATE(rg(k_dim, mesh%gauss%l_G*mesh%me), uex(mesh%gauss%l_G*mesh%me), zero(mesh%np))
    idx = 0
    DO m = 1, mesh%me
       DO l = 1, mesh%gauss%l_G
          idx = idx + 1
          DO k = 1, mesh%gauss%k_d
             r = 0
             DO n = 1, mesh%gauss%n_w
                r = r + mesh%gauss%ww(n,l)*mesh%rr(k,mesh%jj(n,m))
             END DO
             rg(k,idx) = r
          END DO
       END DO
    END DO
    zero = 0.d0
    DO k = 1, 3
       uex = sol_anal(k, rg, inputs%time)
       CALL ns_anal_l1(mesh, un(k,:), uex, out(4*k-3))
       CALL ns_anal_l1(mesh, zero, uex, out(4*k-2))
       CALL ns_anal_0(mesh, un(k,:), uex, out(4*k-1))
       CALL ns_anal_0(mesh, zero, uex, out(4*k))
    END DO
    CALL write_real(arg(6), out)
  END SUBROUTINE do_errors

  !---------------------------------------------------------------------------------------------
  ! The unit-10 diagnostic of the time loop (SOURCES/main.f90:94-99), the same calls and expression.
  SUBROUTINE do_flow
    REAL(KIND=8) :: q1, q2, q3
    REAL(KIND=8), DIMENSION(4) :: out
    CALL prepare(arg(2), arg(3))
    CALL read_state(arg(4), un)
    CALL ns_l1(mesh, un(2,:), q1)
    CALL ns_l1(mesh, un(3,:), q2)
    CALL ns_l1(mesh, un(1,:)*SQRT(un(1,:)), q3)
    out(1) = q1
    out(2) = q2
    out(3) = q3
    out(4) = 0.5d0*(q1+q2)/(inputs%gravity*q3)
    CALL write_real(arg(5), out)
  END SUBROUTINE do_flow

  !---------------------------------------------------------------------------------------------
  ! Fixed-step form of the SSP-RK3 loop of SOURCES/main.f90:61-92.
  SUBROUTINE do_run
    REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE :: ui, uo
    REAL(KIND=8), DIMENSION(:), ALLOCATABLE :: dts
    CHARACTER(LEN=512) :: od, s, fname
    INTEGER :: nsteps, every, it
    INTEGER(KIND=8) :: c0, c1, crate
    REAL(KIND=8) :: to
    CALL prepare(arg(2), arg(3))
    s = arg(4)
    READ(s, *) nsteps
    od = arg(5)
    every = 0
    IF (COMMAND_ARGUMENT_COUNT() >= 6) THEN
       s = arg(6)
       READ(s, *) every
    END IF
    ALLOCATE(ui(inputs%syst_size, mesh%np), uo(inputs%syst_size, mesh%np), dts(MAX(nsteps,1)))
    CALL write_state(TRIM(od)//'/u0.f8', un)
    CALL SYSTEM_CLOCK(c0, crate)
    DO it = 1, nsteps
       IF (inputs%type_test==9) THEN
          inputs%htiny = 1.d-3
          inputs%dt=2.5d-2
       ELSE
          CALL compute_dt_ref(un)
       END IF
       dts(it) = inputs%dt
       to = inputs%time
       uo = un
       CALL euler(uo,un)
       CALL bdy_ref(un,inputs%time+inputs%dt)
       IF (it == 1) CALL write_state(TRIM(od)//'/s1_stage1.f8', un)
       inputs%time=to+inputs%dt
       CALL euler(un,ui)
       un = (3*uo+ ui)/4
       CALL bdy_ref(un,inputs%time+inputs%dt/2)
       IF (it == 1) CALL write_state(TRIM(od)//'/s1_stage2.f8', un)
       inputs%time =  to + inputs%dt/2
       CALL euler(un,ui)
       un = (uo+ 2*ui)/3
       CALL bdy_ref(un,inputs%time+inputs%dt)
       IF (it == 1) CALL write_state(TRIM(od)//'/s1_stage3.f8', un)
       inputs%time = to + inputs%dt
       IF (every > 0) THEN
          IF (MOD(it, every) == 0) THEN
             WRITE(fname, '(A,I6.6,A)') '/snap_', it, '.f8'
             CALL write_state(TRIM(od)//TRIM(fname), un)
          END IF
       END IF
    END DO
    CALL SYSTEM_CLOCK(c1)
    ! wall time of the time loop (the reference's own timer, user_time in SOURCES/main.f90:190-196,
    ! reads the same clock)
    CALL write_real(TRIM(od)//'/loop_seconds.f8', [REAL(c1 - c0, 8) / REAL(crate, 8)])
    CALL write_real(TRIM(od)//'/dt.f8', dts(1:nsteps))
    CALL write_state(TRIM(od)//'/final.f8', un)
    CALL write_real(TRIM(od)//'/time.f8', [inputs%time])
  END SUBROUTINE do_run

END PROGRAM ref_driver
