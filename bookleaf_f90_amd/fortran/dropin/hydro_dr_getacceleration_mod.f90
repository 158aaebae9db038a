! hydro_dr_getacceleration_mod -- drop-in for
! src/packages/hydro/driver/hydro_dr_getacceleration.f90 (:44-112): the
! reference's four-part split of the nodal acceleration, each part on the
! GPU.  The HALFSTEP ghost exchange of hydro_dr_initacceleration has no
! counterpart: one partition per GPU.
MODULE hydro_dr_getacceleration_mod

  USE dataAPI_types_mod,  ONLY: hydro_t,runtime_t,sizes_t,data_t,comm_t
  USE timerAPI_types_mod, ONLY: timer_t
  USE bookleaf_hip_mod,   ONLY: bl_hydro_initacceleration,                    &
&                               bl_hydro_scatteracceleration,                 &
&                               bl_hydro_getacceleration,                     &
&                               bl_hydro_applyacceleration
  USE bookleaf_dropin_mod,ONLY: dropin_attach,dropin_handle,dropin_require,    &
&                               dropin_timestep

  IMPLICIT NONE

  PUBLIC :: hydro_dr_getacceleration,hydro_dr_initacceleration,                &
&           hydro_dr_scatteracceleration,hydro_dr_applyacceleration

CONTAINS

  SUBROUTINE hydro_dr_getacceleration(hydro,sizes,dh)
    TYPE(hydro_t),             INTENT(IN)    :: hydro
    TYPE(sizes_t),             INTENT(IN)    :: sizes
    TYPE(data_t), DIMENSION(:),INTENT(INOUT) :: dh
    CALL dropin_attach(hydro,sizes,dh)
    CALL dropin_require(bl_hydro_getacceleration(dropin_handle()),             &
&                       'hydro_dr_getacceleration')
  END SUBROUTINE hydro_dr_getacceleration

  SUBROUTINE hydro_dr_initacceleration(comm,sizes,timer,dh)
    TYPE(comm_t),              INTENT(IN)    :: comm
    TYPE(sizes_t),             INTENT(IN)    :: sizes
    TYPE(timer_t),             INTENT(INOUT) :: timer
    TYPE(data_t), DIMENSION(:),INTENT(INOUT) :: dh
    CALL dropin_require(bl_hydro_initacceleration(dropin_handle()),            &
&                       'hydro_dr_initacceleration')
  END SUBROUTINE hydro_dr_initacceleration

  SUBROUTINE hydro_dr_scatteracceleration(hydro,sizes,dh)
    TYPE(hydro_t),             INTENT(IN)    :: hydro
    TYPE(sizes_t),             INTENT(IN)    :: sizes
    TYPE(data_t), DIMENSION(:),INTENT(INOUT) :: dh
    CALL dropin_attach(hydro,sizes,dh)
    CALL dropin_require(bl_hydro_scatteracceleration(dropin_handle()),         &
&                       'hydro_dr_scatteracceleration')
  END SUBROUTINE hydro_dr_scatteracceleration

  SUBROUTINE hydro_dr_applyacceleration(runtime,dh)
    TYPE(runtime_t),             INTENT(IN)    :: runtime
    TYPE(data_t),   DIMENSION(:),INTENT(INOUT) :: dh
    CALL dropin_require(bl_hydro_applyacceleration(dropin_handle(),            &
&                                   dropin_timestep(runtime%timestep)),        &
&                       'hydro_dr_applyacceleration')
  END SUBROUTINE hydro_dr_applyacceleration

END MODULE hydro_dr_getacceleration_mod
