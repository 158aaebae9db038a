! hydro_dr_getstate_mod -- drop-in for
! src/packages/hydro/driver/hydro_dr_getstate.f90 (hydro_dr_getstate, :33):
! vertex update by dts, geometry (with its negative-volume error) and
! density on the GPU.
MODULE hydro_dr_getstate_mod

  USE dataAPI_types_mod,  ONLY: runtime_t,data_t,error_t
  USE timerAPI_types_mod, ONLY: timer_t
  USE bookleaf_hip_mod,   ONLY: bl_error_t,bl_hydro_getstate
  USE bookleaf_dropin_mod,ONLY: dropin_handle,dropin_error,dropin_timestep

  IMPLICIT NONE

  PUBLIC :: hydro_dr_getstate

CONTAINS

  SUBROUTINE hydro_dr_getstate(runtime,timer,dh,error)
    TYPE(runtime_t),             INTENT(IN)    :: runtime
    TYPE(timer_t),               INTENT(INOUT) :: timer
    TYPE(data_t),   DIMENSION(:),INTENT(INOUT) :: dh
    TYPE(error_t),               INTENT(OUT)   :: error
    TYPE(bl_error_t) :: berr
    INTEGER          :: rc
    rc=bl_hydro_getstate(dropin_handle(),dropin_timestep(runtime%timestep),berr)
    CALL dropin_error(rc,berr,error)
  END SUBROUTINE hydro_dr_getstate

END MODULE hydro_dr_getstate_mod
