! ale_dr_getmeshstatus_mod -- drop-in for
! src/packages/ale/driver/ale_dr_getmeshstatus.f90 (ale_dr_getmeshstatus,
! :32): the node status of the ALE step on the GPU.
MODULE ale_dr_getmeshstatus_mod

  USE dataAPI_types_mod,  ONLY: ale_t,sizes_t,data_t
  USE timerAPI_types_mod, ONLY: timer_t
  USE bookleaf_hip_mod,   ONLY: bl_ale_getmeshstatus
  USE bookleaf_dropin_mod,ONLY: dropin_ale,dropin_handle,dropin_require

  IMPLICIT NONE

  PUBLIC :: ale_dr_getmeshstatus

CONTAINS

  SUBROUTINE ale_dr_getmeshstatus(ale,sizes,timer,dh)
    TYPE(ale_t),               INTENT(IN)    :: ale
    TYPE(sizes_t),             INTENT(IN)    :: sizes
    TYPE(timer_t),             INTENT(INOUT) :: timer
    TYPE(data_t), DIMENSION(:),INTENT(INOUT) :: dh
    CALL dropin_ale(ale)
    CALL dropin_require(bl_ale_getmeshstatus(dropin_handle()),'ale_dr_getmeshstatus')
  END SUBROUTINE ale_dr_getmeshstatus

END MODULE ale_dr_getmeshstatus_mod
