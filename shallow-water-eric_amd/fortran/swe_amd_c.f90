!==============================================================================================
! swe_amd_c.f90 -- iso_c_binding interface to libswe_amd.so (include/swe_amd.h), the MI355X
! explicit shallow-water step.  The types mirror swe_mesh_desc and swe_params; every routine is a
! BIND(C) interface to the C-ABI entry point of the same name.
!==============================================================================================
MODULE swe_amd_c
  USE iso_c_binding
  IMPLICIT NONE

  INTEGER(c_int), PARAMETER :: SWE_OK = 0, SWE_ENEG_H = 1, SWE_ELAMBDA = 2, SWE_EBADOPT = 3, &
       SWE_EHIP = 4, SWE_EINVAL = 5, SWE_ENOMEM = 6
  INTEGER(c_int), PARAMETER :: SWE_GALERKIN = 0, SWE_VISCOUS = 1, SWE_ENTROPY_VISC = 2, SWE_FCT = 3
  INTEGER(c_int), PARAMETER :: SWE_BC_NONE = 0, SWE_BC_WALLS = 1, SWE_BC_INFLOW = 2, SWE_BC_VALUES = 3, &
       SWE_BC_STAGES = 4

  TYPE, BIND(C) :: swe_params
     REAL(c_double)    :: gravity, htiny, cfl, max_water_h, mannings
     INTEGER(c_int)    :: type_test, viscosity_type, viscous_type, alpha_limit, roundoff_fix
     REAL(c_double)    :: lambda_sgn
  END TYPE swe_params

  TYPE, BIND(C) :: swe_batch_caps
     INTEGER(c_long_long) :: max_nodes, max_entries, max_boundary
     INTEGER(c_int)       :: max_instances
  END TYPE swe_batch_caps

  TYPE, BIND(C) :: swe_mesh_desc
     INTEGER(c_int)    :: np, nnz
     TYPE(c_ptr)       :: ia, ja, cx, cy, lumped, bath
     INTEGER(c_int)    :: n_h_dir, n_ux_dir, n_uy_dir
     TYPE(c_ptr)       :: h_dir, ux_dir, uy_dir
  END TYPE swe_mesh_desc

  INTERFACE
     FUNCTION swe_last_error() BIND(C, NAME='swe_last_error') RESULT(msg)
       IMPORT :: c_ptr
       TYPE(c_ptr) :: msg
     END FUNCTION swe_last_error

     FUNCTION swe_assemble_nnz(np, me, jj) BIND(C, NAME='swe_assemble_nnz') RESULT(nnz)
       IMPORT :: c_int, c_long
       INTEGER(c_int), VALUE :: np, me
       INTEGER(c_int), DIMENSION(*), INTENT(IN) :: jj
       INTEGER(c_long) :: nnz
     END FUNCTION swe_assemble_nnz

     FUNCTION swe_assemble(np, me, jj, rr, ia, ja, diag, mass, lumped, cx, cy) &
          BIND(C, NAME='swe_assemble') RESULT(rc)
       IMPORT :: c_int, c_double
       INTEGER(c_int), VALUE :: np, me
       INTEGER(c_int), DIMENSION(*), INTENT(IN) :: jj
       REAL(c_double), DIMENSION(*), INTENT(IN) :: rr
       INTEGER(c_int), DIMENSION(*) :: ia, ja, diag
       REAL(c_double), DIMENSION(*) :: mass, lumped, cx, cy
       INTEGER(c_int) :: rc
     END FUNCTION swe_assemble

     FUNCTION swe_create(mesh, prm, handle) BIND(C, NAME='swe_create') RESULT(rc)
       IMPORT :: c_int, c_ptr, swe_mesh_desc, swe_params
       TYPE(swe_mesh_desc), INTENT(IN) :: mesh
       TYPE(swe_params), INTENT(IN) :: prm
       TYPE(c_ptr) :: handle
       INTEGER(c_int) :: rc
     END FUNCTION swe_create

     FUNCTION swe_assemble_device(np, me, jj, rr, assembly, nnz) BIND(C, NAME='swe_assemble_device') RESULT(rc)
       IMPORT :: c_int, c_double, c_ptr, c_long_long
       INTEGER(c_int), VALUE :: np, me
       INTEGER(c_int), INTENT(IN) :: jj(*)
       REAL(c_double), INTENT(IN) :: rr(*)
       TYPE(c_ptr) :: assembly
       INTEGER(c_long_long) :: nnz
       INTEGER(c_int) :: rc
     END FUNCTION swe_assemble_device

     FUNCTION swe_assembly_destroy(assembly) BIND(C, NAME='swe_assembly_destroy') RESULT(rc)
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: assembly
       INTEGER(c_int) :: rc
     END FUNCTION swe_assembly_destroy

     FUNCTION swe_create_from_assembly(assembly, bath, n_h_dir, h_dir, n_ux_dir, ux_dir, n_uy_dir, uy_dir, &
          prm, handle) BIND(C, NAME='swe_create_from_assembly') RESULT(rc)
       IMPORT :: c_int, c_double, c_ptr, swe_params
       TYPE(c_ptr), VALUE :: assembly
       REAL(c_double), INTENT(IN) :: bath(*)
       INTEGER(c_int), VALUE :: n_h_dir, n_ux_dir, n_uy_dir
       INTEGER(c_int), INTENT(IN) :: h_dir(*), ux_dir(*), uy_dir(*)
       TYPE(swe_params), INTENT(IN) :: prm
       TYPE(c_ptr) :: handle
       INTEGER(c_int) :: rc
     END FUNCTION swe_create_from_assembly

     FUNCTION swe_errors_create(handle, me, jj, rr, errors) BIND(C, NAME='swe_errors_create') RESULT(rc)
       IMPORT :: c_int, c_double, c_ptr
       TYPE(c_ptr), VALUE :: handle
       INTEGER(c_int), VALUE :: me
       INTEGER(c_int), INTENT(IN) :: jj(*)
       REAL(c_double), INTENT(IN) :: rr(*)
       TYPE(c_ptr) :: errors
       INTEGER(c_int) :: rc
     END FUNCTION swe_errors_create

     FUNCTION swe_compute_errors(errors, time, out) BIND(C, NAME='swe_compute_errors') RESULT(rc)
       IMPORT :: c_int, c_double, c_ptr
       TYPE(c_ptr), VALUE :: errors
       REAL(c_double), VALUE :: time
       REAL(c_double), INTENT(OUT) :: out(12)
       INTEGER(c_int) :: rc
     END FUNCTION swe_compute_errors

     ! the unit-10 diagnostic of tests 1, 2, 10 (main.f90:94-99): q1, q2, q3 and 0.5*(q1+q2)/(g*q3)
     FUNCTION swe_ns_l1_ratio(errors, out) BIND(C, NAME='swe_ns_l1_ratio') RESULT(rc)
       IMPORT :: c_int, c_double, c_ptr
       TYPE(c_ptr), VALUE :: errors
       REAL(c_double), INTENT(OUT) :: out(4)
       INTEGER(c_int) :: rc
     END FUNCTION swe_ns_l1_ratio

     FUNCTION swe_errors_destroy(errors) BIND(C, NAME='swe_errors_destroy') RESULT(rc)
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: errors
       INTEGER(c_int) :: rc
     END FUNCTION swe_errors_destroy

     ! ---- batches: many meshes advanced together (include/swe_amd.h, swe_batch_*) ----
     FUNCTION swe_batch_create(caps, batch) BIND(C, NAME='swe_batch_create') RESULT(rc)
       IMPORT :: c_int, c_ptr, swe_batch_caps
       TYPE(swe_batch_caps), INTENT(IN) :: caps
       TYPE(c_ptr) :: batch
       INTEGER(c_int) :: rc
     END FUNCTION swe_batch_create

     FUNCTION swe_batch_destroy(batch) BIND(C, NAME='swe_batch_destroy') RESULT(rc)
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: batch
       INTEGER(c_int) :: rc
     END FUNCTION swe_batch_destroy

     FUNCTION swe_batch_add(batch, mesh, prm, un, t0, tfinal, max_steps, id) BIND(C, NAME='swe_batch_add') RESULT(rc)
       IMPORT :: c_int, c_ptr, c_double, c_long_long, swe_mesh_desc, swe_params
       TYPE(c_ptr), VALUE :: batch
       TYPE(swe_mesh_desc), INTENT(IN) :: mesh
       TYPE(swe_params), INTENT(IN) :: prm
       REAL(c_double), INTENT(IN) :: un(*)
       REAL(c_double), VALUE :: t0, tfinal
       INTEGER(c_long_long), VALUE :: max_steps
       INTEGER(c_int) :: id
       INTEGER(c_int) :: rc
     END FUNCTION swe_batch_add

     FUNCTION swe_batch_add_assembled(batch, assembly, bath, n_h_dir, h_dir, n_ux_dir, ux_dir, n_uy_dir, uy_dir, &
          prm, un, t0, tfinal, max_steps, id) BIND(C, NAME='swe_batch_add_assembled') RESULT(rc)
       IMPORT :: c_int, c_ptr, c_double, c_long_long, swe_params
       TYPE(c_ptr), VALUE :: batch, assembly
       REAL(c_double), INTENT(IN) :: bath(*), un(*)
       INTEGER(c_int), VALUE :: n_h_dir, n_ux_dir, n_uy_dir
       INTEGER(c_int), INTENT(IN) :: h_dir(*), ux_dir(*), uy_dir(*)
       TYPE(swe_params), INTENT(IN) :: prm
       REAL(c_double), VALUE :: t0, tfinal
       INTEGER(c_long_long), VALUE :: max_steps
       INTEGER(c_int) :: id
       INTEGER(c_int) :: rc
     END FUNCTION swe_batch_add_assembled

     FUNCTION swe_batch_set_bc(batch, id, kind, amp, period, level) BIND(C, NAME='swe_batch_set_bc') RESULT(rc)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: batch
       INTEGER(c_int), VALUE :: id, kind
       REAL(c_double), VALUE :: amp, period, level
       INTEGER(c_int) :: rc
     END FUNCTION swe_batch_set_bc

     FUNCTION swe_batch_remove(batch, id) BIND(C, NAME='swe_batch_remove') RESULT(rc)
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: batch
       INTEGER(c_int), VALUE :: id
       INTEGER(c_int) :: rc
     END FUNCTION swe_batch_remove

     FUNCTION swe_batch_step(batch, nsteps) BIND(C, NAME='swe_batch_step') RESULT(rc)
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: batch
       INTEGER(c_int), VALUE :: nsteps
       INTEGER(c_int) :: rc
     END FUNCTION swe_batch_step

     FUNCTION swe_batch_sync(batch) BIND(C, NAME='swe_batch_sync') RESULT(rc)
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: batch
       INTEGER(c_int) :: rc
     END FUNCTION swe_batch_sync

     FUNCTION swe_batch_status(batch, id, state, time, steps, error) BIND(C, NAME='swe_batch_status') RESULT(rc)
       IMPORT :: c_int, c_ptr, c_double, c_long_long
       TYPE(c_ptr), VALUE :: batch
       INTEGER(c_int), VALUE :: id
       INTEGER(c_int) :: state, error
       REAL(c_double) :: time
       INTEGER(c_long_long) :: steps
       INTEGER(c_int) :: rc
     END FUNCTION swe_batch_status

     FUNCTION swe_batch_get_state(batch, id, un) BIND(C, NAME='swe_batch_get_state') RESULT(rc)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: batch
       INTEGER(c_int), VALUE :: id
       REAL(c_double), INTENT(OUT) :: un(*)
       INTEGER(c_int) :: rc
     END FUNCTION swe_batch_get_state

     FUNCTION swe_batch_errors_create(batch, id, me, jj, rr, errors) BIND(C, NAME='swe_batch_errors_create') &
          RESULT(rc)
       IMPORT :: c_int, c_double, c_ptr
       TYPE(c_ptr), VALUE :: batch
       INTEGER(c_int), VALUE :: id, me
       INTEGER(c_int), INTENT(IN) :: jj(*)
       REAL(c_double), INTENT(IN) :: rr(*)
       TYPE(c_ptr) :: errors
       INTEGER(c_int) :: rc
     END FUNCTION swe_batch_errors_create

     FUNCTION swe_destroy(handle) BIND(C, NAME='swe_destroy') RESULT(rc)
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: handle
       INTEGER(c_int) :: rc
     END FUNCTION swe_destroy

     FUNCTION swe_set_params(handle, prm) BIND(C, NAME='swe_set_params') RESULT(rc)
       IMPORT :: c_int, c_ptr, swe_params
       TYPE(c_ptr), VALUE :: handle
       TYPE(swe_params), INTENT(IN) :: prm
       INTEGER(c_int) :: rc
     END FUNCTION swe_set_params

     FUNCTION swe_compute_dij(handle, un, dij) BIND(C, NAME='swe_compute_dij') RESULT(rc)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: handle
       REAL(c_double), DIMENSION(*), INTENT(IN) :: un
       REAL(c_double), DIMENSION(*) :: dij
       INTEGER(c_int) :: rc
     END FUNCTION swe_compute_dij

     FUNCTION swe_compute_dt(handle, un, dt) BIND(C, NAME='swe_compute_dt') RESULT(rc)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: handle
       REAL(c_double), DIMENSION(*), INTENT(IN) :: un
       REAL(c_double) :: dt
       INTEGER(c_int) :: rc
     END FUNCTION swe_compute_dt

     FUNCTION swe_euler(handle, un, unext, dt) BIND(C, NAME='swe_euler') RESULT(rc)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: handle
       REAL(c_double), DIMENSION(*), INTENT(IN) :: un
       REAL(c_double), DIMENSION(*) :: unext
       REAL(c_double), VALUE :: dt
       INTEGER(c_int) :: rc
     END FUNCTION swe_euler

     ! ---- device-resident time loop (main.f90:71-91 on the GPU) ------------------------------
     FUNCTION swe_set_state(handle, un) BIND(C, NAME='swe_set_state') RESULT(rc)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: handle
       REAL(c_double), DIMENSION(*), INTENT(IN) :: un
       INTEGER(c_int) :: rc
     END FUNCTION swe_set_state

     FUNCTION swe_get_state(handle, un) BIND(C, NAME='swe_get_state') RESULT(rc)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: handle
       REAL(c_double), DIMENSION(*) :: un
       INTEGER(c_int) :: rc
     END FUNCTION swe_get_state

     FUNCTION swe_set_time(handle, t) BIND(C, NAME='swe_set_time') RESULT(rc)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: handle
       REAL(c_double), VALUE :: t
       INTEGER(c_int) :: rc
     END FUNCTION swe_set_time

     FUNCTION swe_get_time(handle, t) BIND(C, NAME='swe_get_time') RESULT(rc)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: handle
       REAL(c_double) :: t
       INTEGER(c_int) :: rc
     END FUNCTION swe_get_time

     FUNCTION swe_set_bc(handle, kind, amp, period, level) BIND(C, NAME='swe_set_bc') RESULT(rc)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: handle
       INTEGER(c_int), VALUE :: kind
       REAL(c_double), VALUE :: amp, period, level
       INTEGER(c_int) :: rc
     END FUNCTION swe_set_bc

     FUNCTION swe_run(handle, nsteps, time, dts) BIND(C, NAME='swe_run') RESULT(rc)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: handle
       INTEGER(c_int), VALUE :: nsteps
       REAL(c_double) :: time
       REAL(c_double), DIMENSION(*) :: dts
       INTEGER(c_int) :: rc
     END FUNCTION swe_run

     FUNCTION swe_step_begin(handle, dt) BIND(C, NAME='swe_step_begin') RESULT(rc)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: handle
       REAL(c_double) :: dt
       INTEGER(c_int) :: rc
     END FUNCTION swe_step_begin

     FUNCTION swe_step_finish(handle, vals) BIND(C, NAME='swe_step_finish') RESULT(rc)
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: handle
       REAL(c_double), DIMENSION(*), INTENT(IN) :: vals
       INTEGER(c_int) :: rc
     END FUNCTION swe_step_finish

     FUNCTION swe_fct_positivity_csr(np, nnz, ia, ja, unext_scaled, f, htiny, lij) &
          BIND(C, NAME='swe_fct_positivity_csr') RESULT(rc)
       IMPORT :: c_int, c_double
       INTEGER(c_int), VALUE :: np, nnz
       INTEGER(c_int), DIMENSION(*), INTENT(IN) :: ia, ja
       REAL(c_double), DIMENSION(*), INTENT(IN) :: unext_scaled, f
       REAL(c_double), VALUE :: htiny
       REAL(c_double), DIMENSION(*) :: lij
       INTEGER(c_int) :: rc
     END FUNCTION swe_fct_positivity_csr
  END INTERFACE

CONTAINS

  ! The library's last error message as a Fortran string.
  FUNCTION swe_error_message() RESULT(s)
    CHARACTER(LEN=512) :: s
    CHARACTER(KIND=c_char), DIMENSION(:), POINTER :: p
    INTEGER :: n
    TYPE(c_ptr) :: m
    s = ' '
    m = swe_last_error()
    IF (.NOT. c_associated(m)) RETURN
    CALL c_f_pointer(m, p, [512])
    DO n = 1, 512
       IF (p(n) == c_null_char) EXIT
       s(n:n) = p(n)
    END DO
  END FUNCTION swe_error_message

END MODULE swe_amd_c
