!==============================================================================================
! fct_swe_amd.f90 -- MODULE fct with the interface of SOURCES/fct.f90:5-42, backed by the MI355X
! library: FCT_positivity(unext, mat, lij) limits on the device and returns lij%aa.
!==============================================================================================
MODULE fct
  USE matrix_type
  USE input_data
  USE iso_c_binding
  USE swe_amd_c
CONTAINS
  SUBROUTINE FCT_positivity(unext,mat,lij)
    IMPLICIT NONE
    TYPE(matrice_bloc),         INTENT(IN)    :: mat
    TYPE(matrice_bloc),         INTENT(INOUT) :: lij
    REAL(KIND=8), DIMENSION(:), INTENT(IN)    :: unext
    INTEGER(c_int) :: rc
    rc = swe_fct_positivity_csr(SIZE(unext), SIZE(mat%ja), mat%ia, mat%ja, unext, mat%aa, inputs%htiny, lij%aa)
    IF (rc /= SWE_OK) THEN
       WRITE(*,*) ' swe_amd error: ', TRIM(swe_error_message())
       STOP
    END IF
  END SUBROUTINE FCT_positivity
END MODULE fct
