!==============================================================================================
! update_swe_amd.f90 -- MODULE update with the public interface of SOURCES/update.f90:1-15
! (construct_matrices, euler, compute_dij and the public arrays dij, lumped, diag, ...), backed by
! the MI355X library through swe_amd_c.  Linking this module instead of SOURCES/update.f90 moves
! the explicit step of the reference driver (SOURCES/main.f90) onto the GPU without changing it.
!
! construct_matrices assembles the P1 matrices with swe_assemble (the reference's element loops,
! bit for bit).  The device solver is created at the first euler or compute_dij call, when init
! has set the bathymetry, gravity and htiny; parameters are refreshed at every call because the
! driver may change them (type_test 9 resets htiny and dt inside the time loop).  Library errors
! end the run with the reference's messages and STOP.
!==============================================================================================
MODULE update
  USE matrix_type
  USE space_dim
  USE input_data
  USE iso_c_binding
  USE swe_amd_c
  PUBLIC :: construct_matrices, euler, compute_dij
  TYPE(matrice_bloc), PUBLIC                  :: dij, betaij
  REAL(KIND=8), DIMENSION(:), POINTER, PUBLIC :: lumped
  INTEGER, DIMENSION(:), POINTER, PUBLIC      :: diag
  REAL(KIND=8), DIMENSION(:), POINTER, PUBLIC :: fix_roundoff
  TYPE(matrice_bloc), PUBLIC                  :: mass, pre_mass, mc_minus_ml
  PRIVATE
  REAL(KIND=8), DIMENSION(:), ALLOCATABLE, TARGET :: cx, cy
  TYPE(c_ptr) :: handle = c_null_ptr
CONTAINS

  SUBROUTINE construct_matrices
    USE mesh_handling
    IMPLICIT NONE
    INTEGER :: np, me, nnz, i, rc
    np = mesh%np
    me = mesh%me
    nnz = INT(swe_assemble_nnz(np, me, mesh%jj))
    IF (nnz <= 0) THEN
       WRITE(*,*) ' BUG in construct_matrices: invalid mesh'
       STOP
    END IF
    ALLOCATE(mass%ia(np+1), mass%ja(nnz), mass%aa(nnz), lumped(np), diag(np), cx(nnz), cy(nnz))
    rc = swe_assemble(np, me, mesh%jj, mesh%rr, mass%ia, mass%ja, diag, mass%aa, lumped, cx, cy)
    CALL check(rc)
    ALLOCATE(fix_roundoff(np))
    fix_roundoff = 0.d0
    !===Mass - lumped and preconditioned mass, as in SOURCES/update.f90:52-62
    mc_minus_ml%ia => mass%ia
    mc_minus_ml%ja => mass%ja
    ALLOCATE(mc_minus_ml%aa(nnz))
    mc_minus_ml%aa = mass%aa
    mc_minus_ml%aa(diag) = mc_minus_ml%aa(diag) - lumped
    pre_mass%ia => mass%ia
    pre_mass%ja => mass%ja
    ALLOCATE(pre_mass%aa(nnz))
    DO i = 1, np
       pre_mass%aa(mass%ia(i):mass%ia(i+1)-1) = mass%aa(mass%ia(i):mass%ia(i+1)-1)/lumped(i)
    END DO
    !===dij shares the pattern
    dij%ia => mass%ia
    dij%ja => mass%ja
    ALLOCATE(dij%aa(nnz))
    dij%aa = 0.d0
  END SUBROUTINE construct_matrices

  SUBROUTINE euler(un,unext)
    USE mesh_handling
    IMPLICIT NONE
    REAL(KIND=8), DIMENSION(inputs%syst_size,mesh%np), INTENT(IN)  :: un
    REAL(KIND=8), DIMENSION(inputs%syst_size,mesh%np), INTENT(OUT) :: unext
    CALL ensure_handle
    CALL check(swe_euler(handle, un, unext, inputs%dt))
  END SUBROUTINE euler

  SUBROUTINE compute_dij(un)
    USE mesh_handling
    IMPLICIT NONE
    REAL(KIND=8), DIMENSION(inputs%syst_size,mesh%np), INTENT(IN)  :: un
    CALL ensure_handle
    CALL check(swe_compute_dij(handle, un, dij%aa))
  END SUBROUTINE compute_dij

  FUNCTION current_params() RESULT(prm)
    USE boundary_conditions
    IMPLICIT NONE
    TYPE(swe_params) :: prm
    prm%gravity = inputs%gravity
    prm%htiny = inputs%htiny
    prm%cfl = inputs%CFL
    prm%max_water_h = max_water_h
    prm%mannings = inputs%mannings
    prm%type_test = inputs%type_test
    SELECT CASE (TRIM(ADJUSTL(inputs%viscosity_type)))
    CASE ('galerkin')
       prm%viscosity_type = SWE_GALERKIN
    CASE ('viscous')
       prm%viscosity_type = SWE_VISCOUS
    CASE ('entropy_visc')
       prm%viscosity_type = SWE_ENTROPY_VISC
    CASE ('fct')
       prm%viscosity_type = SWE_FCT
    CASE DEFAULT
       prm%viscosity_type = -1
    END SELECT
    SELECT CASE (TRIM(ADJUSTL(inputs%viscous_type)))
    CASE ('type1')
       prm%viscous_type = 1
    CASE ('type2')
       prm%viscous_type = 2
    CASE DEFAULT
       prm%viscous_type = -1
    END SELECT
    prm%alpha_limit = MERGE(1, 0, inputs%if_alpha_limit)
    prm%roundoff_fix = 1
    prm%lambda_sgn = 0.d0
    IF (inputs%type_test == 13) prm%lambda_sgn = inputs%lambdaSGN
  END FUNCTION current_params

  SUBROUTINE ensure_handle
    USE mesh_handling
    USE boundary_conditions
    IMPLICIT NONE
    TYPE(swe_mesh_desc) :: desc
    TYPE(swe_params) :: prm
    REAL(KIND=8), DIMENSION(:), ALLOCATABLE, TARGET :: z
    IF (inputs%syst_size /= MERGE(5, 3, inputs%type_test == 13)) THEN
       WRITE(*,*) ' BUG: the GPU step is programmed for syst_size = 3 (5 for the SGN model)'
       STOP
    END IF
    IF (.NOT. inputs%if_lumped) THEN
       WRITE(*,*) ' BUG: the GPU step is programmed with the lumped mass only'
       STOP
    END IF
    prm = current_params()
    IF (c_associated(handle)) THEN
       CALL check(swe_set_params(handle, prm))
       RETURN
    END IF
    desc%np = mesh%np
    desc%nnz = SIZE(mass%ja)
    desc%ia = c_loc(mass%ia(1))
    desc%ja = c_loc(mass%ja(1))
    desc%cx = c_loc(cx(1))
    desc%cy = c_loc(cy(1))
    desc%lumped = c_loc(lumped(1))
    ALLOCATE(z(mesh%np))
    z = bath
    desc%bath = c_loc(z(1))
    desc%n_h_dir = 0
    desc%n_ux_dir = 0
    desc%n_uy_dir = 0
    desc%h_dir = c_null_ptr
    desc%ux_dir = c_null_ptr
    desc%uy_dir = c_null_ptr
    CALL check(swe_create(desc, prm, handle))
  END SUBROUTINE ensure_handle

  ! Library status -> the reference's stop messages (SOURCES/update.f90:199,217,817; condlim.f90:935)
  SUBROUTINE check(rc)
    IMPLICIT NONE
    INTEGER(c_int), INTENT(IN) :: rc
    SELECT CASE (rc)
    CASE (SWE_OK)
       RETURN
    CASE (SWE_ENEG_H)
       WRITE(*,*) 'Min h<0, STOP'
    CASE (SWE_ELAMBDA)
       WRITE(*,*) ' BUG in compute_lambda_vacc'
    CASE (SWE_EBADOPT)
       WRITE(*,*) ' BUG in euler: ', TRIM(swe_error_message())
    CASE DEFAULT
       WRITE(*,*) ' swe_amd error: ', TRIM(swe_error_message())
    END SELECT
    STOP
  END SUBROUTINE check

END MODULE update
