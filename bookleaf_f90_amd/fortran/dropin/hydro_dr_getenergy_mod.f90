! hydro_dr_getenergy_mod -- drop-in for
! src/packages/hydro/driver/hydro_dr_getenergy.f90 (hydro_dr_getenergy,
! :35): the internal energy update (and the components') on the GPU.
MODULE hydro_dr_getenergy_mod

  USE dataAPI_types_mod,  ONLY: hydro_t,runtime_t,data_t
  USE bookleaf_hip_mod,   ONLY: bl_hydro_getenergy
  USE bookleaf_dropin_mod,ONLY: dropin_attach,dropin_handle,dropin_require,    &
&                               dropin_timestep

  IMPLICIT NONE

  PUBLIC :: hydro_dr_getenergy

CONTAINS

  SUBROUTINE hydro_dr_getenergy(hydro,runtime,dh)
    TYPE(hydro_t),               INTENT(IN)    :: hydro
    TYPE(runtime_t),             INTENT(IN)    :: runtime
    TYPE(data_t),   DIMENSION(:),INTENT(INOUT) :: dh
    CALL dropin_attach(hydro,runtime%sizes,dh)
    CALL dropin_require(bl_hydro_getenergy(dropin_handle(),                    &
&                                          dropin_timestep(runtime%timestep)), &
&                       'hydro_dr_getenergy')
  END SUBROUTINE hydro_dr_getenergy

END MODULE hydro_dr_getenergy_mod
