! hydro_dr_getdt_mod -- drop-in for src/packages/hydro/driver/hydro_dr_getdt.f90
! (hydro_dr_getdt, :39): the viscosity, CFL and divergence controls of the
! step on the GPU (bl_hydro_getdt), appended to the dt_t list as the
! reference appends them, CFL then DIV.  The first call attaches the GPU
! handle (bookleaf_dropin_mod).
MODULE hydro_dr_getdt_mod

  USE dataAPI_types_mod,  ONLY: hydro_t,sizes_t,error_t,dt_t,data_t
  USE dataAPI_params_mod, ONLY: SUCCESS
  USE timerAPI_types_mod, ONLY: timer_t
  USE bookleaf_hip_mod,   ONLY: bl_dt_t,bl_error_t,bl_hydro_getdt,BL_SUCCESS
  USE bookleaf_dropin_mod,ONLY: dropin_attach,dropin_handle,dropin_error

  IMPLICIT NONE

  PUBLIC :: hydro_dr_getdt

CONTAINS

  SUBROUTINE hydro_dr_getdt(hydro,sizes,timer,dh,dt,error)

    TYPE(hydro_t),             INTENT(IN)    :: hydro
    TYPE(sizes_t),             INTENT(IN)    :: sizes
    TYPE(timer_t),DIMENSION(:),INTENT(INOUT) :: timer
    TYPE(data_t), DIMENSION(:),INTENT(INOUT) :: dh
    TYPE(dt_t),   POINTER,     INTENT(INOUT) :: dt
    TYPE(error_t),             INTENT(OUT)   :: error
    TYPE(bl_dt_t)    :: cfl,div
    TYPE(bl_error_t) :: berr
    INTEGER          :: rc

    CALL dropin_attach(hydro,sizes,dh)
    rc=bl_hydro_getdt(dropin_handle(),cfl,div,berr)
    CALL dropin_error(rc,berr,error)
    ! hydro_kn_getdtcfl's candidate (a negative CFL returns after it)
    ALLOCATE(dt%next)
    dt=>dt%next
    dt%rdt=cfl%rdt
    dt%idt=cfl%idt
    dt%sdt='     CFL'
    IF (error%ierr.NE.SUCCESS) RETURN
    ! hydro_kn_getdtdiv's candidate
    ALLOCATE(dt%next)
    dt=>dt%next
    dt%rdt=div%rdt
    dt%idt=div%idt
    dt%sdt='     DIV'

  END SUBROUTINE hydro_dr_getdt

END MODULE hydro_dr_getdt_mod
