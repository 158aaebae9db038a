! hydro_dr_geteos_mod -- drop-in for
! src/packages/hydro/driver/hydro_dr_geteos.f90 (hydro_dr_geteos, :31):
! pressure and sound speed of elements and components on the GPU.
MODULE hydro_dr_geteos_mod

  USE dataAPI_types_mod,  ONLY: hydro_t,sizes_t,data_t
  USE timerAPI_types_mod, ONLY: timer_t
  USE bookleaf_hip_mod,   ONLY: bl_hydro_geteos
  USE bookleaf_dropin_mod,ONLY: dropin_attach,dropin_handle,dropin_require

  IMPLICIT NONE

  PUBLIC :: hydro_dr_geteos

CONTAINS

  SUBROUTINE hydro_dr_geteos(hydro,sizes,timer,dh)
    TYPE(hydro_t),             INTENT(IN)    :: hydro
    TYPE(sizes_t),             INTENT(IN)    :: sizes
    TYPE(timer_t),DIMENSION(:),INTENT(INOUT) :: timer
    TYPE(data_t), DIMENSION(:),INTENT(INOUT) :: dh
    CALL dropin_attach(hydro,sizes,dh)
    CALL dropin_require(bl_hydro_geteos(dropin_handle()),'hydro_dr_geteos')
  END SUBROUTINE hydro_dr_geteos

END MODULE hydro_dr_geteos_mod
