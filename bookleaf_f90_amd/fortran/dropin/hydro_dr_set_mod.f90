! hydro_dr_set_mod -- drop-in for src/packages/hydro/driver/hydro_dr_set.f90
! (hydro_dr_setpredictor :31, hydro_dr_setcorrector :42): the energy copy
! between elenergy and rscratch11 on the GPU.
MODULE hydro_dr_set_mod

  USE dataAPI_types_mod,  ONLY: sizes_t,data_t
  USE bookleaf_hip_mod,   ONLY: bl_hydro_setpredictor,bl_hydro_setcorrector
  USE bookleaf_dropin_mod,ONLY: dropin_handle,dropin_require

  IMPLICIT NONE

  PUBLIC :: hydro_dr_setpredictor,hydro_dr_setcorrector

CONTAINS

  SUBROUTINE hydro_dr_setpredictor(sizes,dh)
    TYPE(sizes_t),             INTENT(IN)    :: sizes
    TYPE(data_t), DIMENSION(:),INTENT(INOUT) :: dh
    CALL dropin_require(bl_hydro_setpredictor(dropin_handle()),'hydro_dr_setpredictor')
  END SUBROUTINE hydro_dr_setpredictor

  SUBROUTINE hydro_dr_setcorrector(sizes,dh)
    TYPE(sizes_t),             INTENT(IN)    :: sizes
    TYPE(data_t), DIMENSION(:),INTENT(INOUT) :: dh
    CALL dropin_require(bl_hydro_setcorrector(dropin_handle()),'hydro_dr_setcorrector')
  END SUBROUTINE hydro_dr_setcorrector

END MODULE hydro_dr_set_mod
