! ale_dr_getfluxvolume_mod -- drop-in for
! src/packages/ale/driver/ale_dr_getfluxvolume.f90 (ale_dr_getfluxvolume,
! :37): mesh velocity, mesh motion and face flux volumes on the GPU.
MODULE ale_dr_getfluxvolume_mod

  USE dataAPI_types_mod,  ONLY: ale_t,runtime_t,data_t
  USE timerAPI_types_mod, ONLY: timer_t
  USE bookleaf_hip_mod,   ONLY: bl_ale_getfluxvolume
  USE bookleaf_dropin_mod,ONLY: dropin_ale,dropin_handle,dropin_require,       &
&                               dropin_timestep

  IMPLICIT NONE

  PUBLIC :: ale_dr_getfluxvolume

CONTAINS

  SUBROUTINE ale_dr_getfluxvolume(ale,runtime,timer,dh)
    TYPE(ale_t),                 INTENT(IN)    :: ale
    TYPE(runtime_t),             INTENT(IN)    :: runtime
    TYPE(timer_t),               INTENT(INOUT) :: timer
    TYPE(data_t),   DIMENSION(:),INTENT(INOUT) :: dh
    CALL dropin_ale(ale)
    CALL dropin_require(bl_ale_getfluxvolume(dropin_handle(),                  &
&                                            dropin_timestep(runtime%timestep)),&
&                       'ale_dr_getfluxvolume')
  END SUBROUTINE ale_dr_getfluxvolume

END MODULE ale_dr_getfluxvolume_mod
