! ale_dr_advect_mod -- drop-in for src/packages/ale/driver/ale_dr_advect.f90
! (ale_dr_advect, :56): the element and nodal remap, unsplit or split
! (ale%adv_type), on the GPU.  The ADVEL/ADVND ghost exchanges have no
! counterpart: one partition per GPU.
MODULE ale_dr_advect_mod

  USE dataAPI_types_mod,  ONLY: ale_t,runtime_t,error_t,data_t
  USE timerAPI_types_mod, ONLY: timer_t
  USE bookleaf_hip_mod,   ONLY: bl_error_t,bl_ale_advect
  USE bookleaf_dropin_mod,ONLY: dropin_ale,dropin_handle,dropin_error,         &
&                               dropin_timestep

  IMPLICIT NONE

  PUBLIC :: ale_dr_advect

CONTAINS

  SUBROUTINE ale_dr_advect(ale,runtime,timer,dh,error)
    TYPE(ale_t),                 INTENT(IN)    :: ale
    TYPE(runtime_t),             INTENT(IN)    :: runtime
    TYPE(timer_t),  DIMENSION(:),INTENT(INOUT) :: timer
    TYPE(data_t),   DIMENSION(:),INTENT(INOUT) :: dh
    TYPE(error_t),               INTENT(OUT)   :: error
    TYPE(bl_error_t) :: berr
    INTEGER          :: rc
    CALL dropin_ale(ale)
    rc=bl_ale_advect(dropin_handle(),dropin_timestep(runtime%timestep),berr)
    CALL dropin_error(rc,berr,error)
  END SUBROUTINE ale_dr_advect

END MODULE ale_dr_advect_mod
