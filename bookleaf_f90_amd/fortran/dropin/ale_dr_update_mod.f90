! ale_dr_update_mod -- drop-in for src/packages/ale/driver/ale_dr_update.f90
! (ale_dr_hydroupdate, :33): geometry, density and EOS on the remapped mesh
! on the GPU, with the geometry's negative-volume error.
MODULE ale_dr_update_mod

  USE dataAPI_types_mod,  ONLY: ale_t,sizes_t,data_t,error_t
  USE timerAPI_types_mod, ONLY: timer_t
  USE bookleaf_hip_mod,   ONLY: bl_error_t,bl_ale_hydroupdate
  USE bookleaf_dropin_mod,ONLY: dropin_ale,dropin_handle,dropin_error

  IMPLICIT NONE

  PUBLIC :: ale_dr_hydroupdate

CONTAINS

  SUBROUTINE ale_dr_hydroupdate(ale,sizes,timer,dh,error)
    TYPE(ale_t),               INTENT(IN)    :: ale
    TYPE(sizes_t),             INTENT(IN)    :: sizes
    TYPE(timer_t),DIMENSION(:),INTENT(INOUT) :: timer
    TYPE(data_t), DIMENSION(:),INTENT(INOUT) :: dh
    TYPE(error_t),             INTENT(OUT)   :: error
    TYPE(bl_error_t) :: berr
    INTEGER          :: rc
    CALL dropin_ale(ale)
    rc=bl_ale_hydroupdate(dropin_handle(),berr)
    CALL dropin_error(rc,berr,error)
  END SUBROUTINE ale_dr_hydroupdate

END MODULE ale_dr_update_mod
