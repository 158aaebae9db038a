! hydro_dr_getforce_mod -- drop-in for
! src/packages/hydro/driver/hydro_dr_getforce.f90 (hydro_dr_getforce, :40):
! the corner forces of the predictor or corrector on the GPU.
MODULE hydro_dr_getforce_mod

  USE dataAPI_types_mod,  ONLY: hydro_t,runtime_t,data_t
  USE timerAPI_types_mod, ONLY: timer_t
  USE bookleaf_hip_mod,   ONLY: bl_hydro_getforce
  USE bookleaf_dropin_mod,ONLY: dropin_attach,dropin_handle,dropin_require,    &
&                               dropin_timestep

  IMPLICIT NONE

  PUBLIC :: hydro_dr_getforce

CONTAINS

  SUBROUTINE hydro_dr_getforce(hydro,runtime,timer,dh)
    TYPE(hydro_t),               INTENT(IN)    :: hydro
    TYPE(runtime_t),             INTENT(IN)    :: runtime
    TYPE(timer_t),  DIMENSION(:),INTENT(INOUT) :: timer
    TYPE(data_t),   DIMENSION(:),INTENT(INOUT) :: dh
    CALL dropin_attach(hydro,runtime%sizes,dh)
    CALL dropin_require(bl_hydro_getforce(dropin_handle(),                     &
&                                         dropin_timestep(runtime%timestep)),  &
&                       'hydro_dr_getforce')
  END SUBROUTINE hydro_dr_getforce

END MODULE hydro_dr_getforce_mod
