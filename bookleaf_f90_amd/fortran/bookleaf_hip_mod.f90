! bookleaf_hip_mod.f90 -- Fortran binding of libbookleaf_hip.so (C-ABI of
! include/bookleaf_hip.h) for a BookLeaf host.
!
! Every entry point of the C header has a BIND(C) interface here with the
! same name.  The derived types match the C structs field for field.  Field
! ids are the reference's own dataAPI ids (dataAPI_id.f90:26-74: eldensityid
! = 1 ... rcpscratch24id = 79), so a host passes its dh(id) arrays as
!   bl_upload(h, id, C_LOC(dh(id)%raddr(1)), INT(dh(id)%dsize, C_INT64_T)).
! All functions return BL_SUCCESS (0) or BL_FAILURE (-1)
! (dataAPI_params.f90:31-34).
MODULE bookleaf_hip_mod

  USE, INTRINSIC :: iso_c_binding, ONLY: C_INT32_T,C_INT64_T,C_DOUBLE,C_FLOAT,C_CHAR,  &
&                                        C_PTR,C_NULL_CHAR,C_ASSOCIATED,       &
&                                        C_F_POINTER

  IMPLICIT NONE

  INTEGER(KIND=C_INT32_T),PARAMETER :: BL_SUCCESS=0,BL_FAILURE=-1,            &
&                                      BL_HALT_ALL=1,BL_HALT_SINGLE=2,        &
&                                      BL_SLEN=130
  INTEGER(KIND=C_INT32_T),PARAMETER :: BL_DT_MAXIMUM=1,BL_DT_GROWTH=2,        &
&                                      BL_DT_INITIAL=3,BL_DT_CFL=4,BL_DT_DIV=5
  ! node relaxation types (setup_mesh's region types)
  INTEGER(KIND=C_INT32_T),PARAMETER :: BL_RELAX_OFF=0,BL_RELAX_LIN1=1,         &
&                                      BL_RELAX_LIN2=2,BL_RELAX_EQUI=3
  INTEGER(KIND=C_INT32_T),PARAMETER :: BL_KIND_R64=1,BL_KIND_I32=2,           &
&                                      BL_KIND_BOOL=3

  ! sizes_t (dataAPI_types.f90:41-44)
  TYPE,BIND(C) :: bl_sizes_t
    INTEGER(KIND=C_INT32_T) :: nel,nel1,nel2,nnd,nnd1,nnd2,nsz,nmat,nreg,ncp,  &
&                              nmx,mcp,mmx
  END TYPE bl_sizes_t

  ! global_t, hydro_t, eos_t, time_t and ale_t scalars
  TYPE,BIND(C) :: bl_config_t
    REAL(KIND=C_DOUBLE)     :: zcut,zerocut,dencut,accut
    REAL(KIND=C_DOUBLE)     :: cvisc1,cvisc2,cfl_sf,div_sf
    INTEGER(KIND=C_INT32_T) :: zhg,zsp,ztq,pad0
    REAL(KIND=C_DOUBLE)     :: ccut,pcut
    REAL(KIND=C_DOUBLE)     :: time_start,time_end,dt_g,dt_min,dt_max,dt_initial
    REAL(KIND=C_DOUBLE)     :: mintime,maxtime,sf
    INTEGER(KIND=C_INT32_T) :: adv_type,zexist,zon,zeul
  END TYPE bl_config_t

  ! timestep_t (dataAPI_types.f90:46-52), sdt as a BL_DT_* code
  TYPE,BIND(C) :: bl_timestep_t
    REAL(KIND=C_DOUBLE)     :: time,dt,dts
    INTEGER(KIND=C_INT32_T) :: nstep,idtel,zcorrector,sdt
  END TYPE bl_timestep_t

  ! dt_t (dataAPI_types.f90:54-60) without the list pointer
  TYPE,BIND(C) :: bl_dt_t
    REAL(KIND=C_DOUBLE)     :: rdt
    INTEGER(KIND=C_INT32_T) :: idt,sdt
  END TYPE bl_dt_t

  ! error_t (dataAPI_types.f90:76-79)
  TYPE,BIND(C) :: bl_error_t
    INTEGER(KIND=C_INT32_T)                     :: ierr,iout
    CHARACTER(KIND=C_CHAR),DIMENSION(BL_SLEN)   :: serr
  END TYPE bl_error_t

  ! conservation totals (hydro_kn_print.F90:56-110)
  TYPE,BIND(C) :: bl_totals_t
    REAL(KIND=C_DOUBLE) :: volume,mass,ie,ke,pressure_min,pressure_max,        &
&                          density_min,density_max
  END TYPE bl_totals_t

  ! one instance of an ensemble (bl_ensemble_create)
  TYPE,BIND(C) :: bl_instance_t
    INTEGER(KIND=C_INT32_T) :: nx,ny,kind,pad
    REAL(KIND=C_DOUBLE)     :: length,height,gamma,kappa,rho_l,e_l,rho_r,e_r
  END TYPE bl_instance_t

  ! one row of hydro_dr_shortprint's region or material table
  ! (hydro_kn_print.F90:56-110, hydro_dr_print.f90:49-119)
  TYPE,BIND(C) :: bl_flag_totals_t
    REAL(KIND=C_DOUBLE) :: volume,mass,ie,ke,pressure,pressure_min,            &
&                          pressure_max,density_min,density_max
  END TYPE bl_flag_totals_t

  ! the launch shape bl_autotune chose (no reference counterpart)
  TYPE,BIND(C) :: bl_tuning_t
    INTEGER(KIND=C_INT32_T) :: limiter_block,limiter_grid,ens_chunk,trials
    REAL(KIND=C_DOUBLE)     :: best_us,default_us,best_full_us,default_full_us
  END TYPE bl_tuning_t

  INTERFACE

    ! ---- lifetime and registry -------------------------------------------
    INTEGER(KIND=C_INT32_T) FUNCTION bl_create(sizes,h) BIND(C,NAME="bl_create")
      IMPORT :: C_INT32_T,C_PTR,bl_sizes_t
      TYPE(bl_sizes_t),INTENT(IN)  :: sizes
      TYPE(C_PTR),     INTENT(OUT) :: h
    END FUNCTION bl_create
    INTEGER(KIND=C_INT32_T) FUNCTION bl_destroy(h) BIND(C,NAME="bl_destroy")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: h
    END FUNCTION bl_destroy
    INTEGER(KIND=C_INT32_T) FUNCTION bl_set_config(h,cfg) BIND(C,NAME="bl_set_config")
      IMPORT :: C_INT32_T,C_PTR,bl_config_t
      TYPE(C_PTR),VALUE            :: h
      TYPE(bl_config_t),INTENT(IN) :: cfg
    END FUNCTION bl_set_config
    INTEGER(KIND=C_INT32_T) FUNCTION bl_set_regions(h,nreg,kappareg,pmeritreg,  &
&                                                   zdtnotreg,zmidlength)       &
&                                                   BIND(C,NAME="bl_set_regions")
      IMPORT :: C_INT32_T,C_PTR,C_DOUBLE
      TYPE(C_PTR),VALUE                                :: h
      INTEGER(KIND=C_INT32_T),VALUE                    :: nreg
      REAL(KIND=C_DOUBLE),    DIMENSION(*),INTENT(IN)  :: kappareg,pmeritreg
      INTEGER(KIND=C_INT32_T),DIMENSION(*),INTENT(IN)  :: zdtnotreg,zmidlength
    END FUNCTION bl_set_regions
    INTEGER(KIND=C_INT32_T) FUNCTION bl_set_eos(h,nmat,itype,param)             &
&                                               BIND(C,NAME="bl_set_eos")
      IMPORT :: C_INT32_T,C_PTR,C_DOUBLE
      TYPE(C_PTR),VALUE                               :: h
      INTEGER(KIND=C_INT32_T),VALUE                   :: nmat
      INTEGER(KIND=C_INT32_T),DIMENSION(*),INTENT(IN) :: itype
      REAL(KIND=C_DOUBLE),    DIMENSION(*),INTENT(IN) :: param
    END FUNCTION bl_set_eos
    INTEGER(KIND=C_INT32_T) FUNCTION bl_field_info(h,id,kind,count)             &
&                                                  BIND(C,NAME="bl_field_info")
      IMPORT :: C_INT32_T,C_INT64_T,C_PTR
      TYPE(C_PTR),VALUE                    :: h
      INTEGER(KIND=C_INT32_T),VALUE        :: id
      INTEGER(KIND=C_INT32_T),INTENT(OUT)  :: kind
      INTEGER(KIND=C_INT64_T),INTENT(OUT)  :: count
    END FUNCTION bl_field_info
    TYPE(C_PTR) FUNCTION bl_field_device_ptr(h,id) BIND(C,NAME="bl_field_device_ptr")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE             :: h
      INTEGER(KIND=C_INT32_T),VALUE :: id
    END FUNCTION bl_field_device_ptr
    INTEGER(KIND=C_INT32_T) FUNCTION bl_upload(h,id,host,count) BIND(C,NAME="bl_upload")
      IMPORT :: C_INT32_T,C_INT64_T,C_PTR
      TYPE(C_PTR),VALUE             :: h,host
      INTEGER(KIND=C_INT32_T),VALUE :: id
      INTEGER(KIND=C_INT64_T),VALUE :: count
    END FUNCTION bl_upload
    INTEGER(KIND=C_INT32_T) FUNCTION bl_download(h,id,host,count) BIND(C,NAME="bl_download")
      IMPORT :: C_INT32_T,C_INT64_T,C_PTR
      TYPE(C_PTR),VALUE             :: h,host
      INTEGER(KIND=C_INT32_T),VALUE :: id
      INTEGER(KIND=C_INT64_T),VALUE :: count
    END FUNCTION bl_download
    INTEGER(KIND=C_INT32_T) FUNCTION bl_build_maps(h) BIND(C,NAME="bl_build_maps")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: h
    END FUNCTION bl_build_maps
    INTEGER(KIND=C_INT32_T) FUNCTION bl_init_mesh(h) BIND(C,NAME="bl_init_mesh")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: h
    END FUNCTION bl_init_mesh
    INTEGER(KIND=C_INT32_T) FUNCTION bl_synchronize(h) BIND(C,NAME="bl_synchronize")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: h
    END FUNCTION bl_synchronize

    ! ---- initialisation drivers ------------------------------------------
    INTEGER(KIND=C_INT32_T) FUNCTION bl_init_state(h) BIND(C,NAME="bl_init_state")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: h
    END FUNCTION bl_init_state
    INTEGER(KIND=C_INT32_T) FUNCTION bl_hydro_init(h) BIND(C,NAME="bl_hydro_init")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: h
    END FUNCTION bl_hydro_init

    ! ---- Lagrangian step drivers -----------------------------------------
    INTEGER(KIND=C_INT32_T) FUNCTION bl_hydro_getdt(h,cfl,div,err)              &
&                                                   BIND(C,NAME="bl_hydro_getdt")
      IMPORT :: C_INT32_T,C_PTR,bl_dt_t,bl_error_t
      TYPE(C_PTR),VALUE              :: h
      TYPE(bl_dt_t),   INTENT(OUT)   :: cfl,div
      TYPE(bl_error_t),INTENT(OUT)   :: err
    END FUNCTION bl_hydro_getdt
    ! the same with no device allocation in the call (meshes of at most
    ! 2**24 elements; calls that overlap on one GPU race)
    INTEGER(KIND=C_INT32_T) FUNCTION bl_hydro_getdt_noalloc(h,cfl,div,err)      &
&                                                   BIND(C,NAME="bl_hydro_getdt_noalloc")
      IMPORT :: C_INT32_T,C_PTR,bl_dt_t,bl_error_t
      TYPE(C_PTR),VALUE              :: h
      TYPE(bl_dt_t),   INTENT(OUT)   :: cfl,div
      TYPE(bl_error_t),INTENT(OUT)   :: err
    END FUNCTION bl_hydro_getdt_noalloc
    INTEGER(KIND=C_INT32_T) FUNCTION bl_hydro_setpredictor(h)                   &
&                                                   BIND(C,NAME="bl_hydro_setpredictor")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: h
    END FUNCTION bl_hydro_setpredictor
    INTEGER(KIND=C_INT32_T) FUNCTION bl_hydro_setcorrector(h)                   &
&                                                   BIND(C,NAME="bl_hydro_setcorrector")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: h
    END FUNCTION bl_hydro_setcorrector
    INTEGER(KIND=C_INT32_T) FUNCTION bl_hydro_getforce(h,ts) BIND(C,NAME="bl_hydro_getforce")
      IMPORT :: C_INT32_T,C_PTR,bl_timestep_t
      TYPE(C_PTR),VALUE              :: h
      TYPE(bl_timestep_t),INTENT(IN) :: ts
    END FUNCTION bl_hydro_getforce
    INTEGER(KIND=C_INT32_T) FUNCTION bl_hydro_getstate(h,ts,err)                &
&                                                   BIND(C,NAME="bl_hydro_getstate")
      IMPORT :: C_INT32_T,C_PTR,bl_timestep_t,bl_error_t
      TYPE(C_PTR),VALUE               :: h
      TYPE(bl_timestep_t),INTENT(IN)  :: ts
      TYPE(bl_error_t),   INTENT(OUT) :: err
    END FUNCTION bl_hydro_getstate
    INTEGER(KIND=C_INT32_T) FUNCTION bl_hydro_getenergy(h,ts) BIND(C,NAME="bl_hydro_getenergy")
      IMPORT :: C_INT32_T,C_PTR,bl_timestep_t
      TYPE(C_PTR),VALUE              :: h
      TYPE(bl_timestep_t),INTENT(IN) :: ts
    END FUNCTION bl_hydro_getenergy
    INTEGER(KIND=C_INT32_T) FUNCTION bl_hydro_geteos(h) BIND(C,NAME="bl_hydro_geteos")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: h
    END FUNCTION bl_hydro_geteos
    INTEGER(KIND=C_INT32_T) FUNCTION bl_hydro_initacceleration(h)               &
&                                          BIND(C,NAME="bl_hydro_initacceleration")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: h
    END FUNCTION bl_hydro_initacceleration
    INTEGER(KIND=C_INT32_T) FUNCTION bl_hydro_scatteracceleration(h)            &
&                                          BIND(C,NAME="bl_hydro_scatteracceleration")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: h
    END FUNCTION bl_hydro_scatteracceleration
    INTEGER(KIND=C_INT32_T) FUNCTION bl_hydro_getacceleration(h)                &
&                                          BIND(C,NAME="bl_hydro_getacceleration")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: h
    END FUNCTION bl_hydro_getacceleration
    INTEGER(KIND=C_INT32_T) FUNCTION bl_hydro_applyacceleration(h,ts)           &
&                                          BIND(C,NAME="bl_hydro_applyacceleration")
      IMPORT :: C_INT32_T,C_PTR,bl_timestep_t
      TYPE(C_PTR),VALUE              :: h
      TYPE(bl_timestep_t),INTENT(IN) :: ts
    END FUNCTION bl_hydro_applyacceleration

    ! ---- shared utilities ------------------------------------------------
    INTEGER(KIND=C_INT32_T) FUNCTION bl_geometry_getgeometry(h,err)             &
&                                          BIND(C,NAME="bl_geometry_getgeometry")
      IMPORT :: C_INT32_T,C_PTR,bl_error_t
      TYPE(C_PTR),VALUE             :: h
      TYPE(bl_error_t),INTENT(OUT)  :: err
    END FUNCTION bl_geometry_getgeometry
    INTEGER(KIND=C_INT32_T) FUNCTION bl_density_getdensity(h)                   &
&                                          BIND(C,NAME="bl_density_getdensity")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: h
    END FUNCTION bl_density_getdensity

    ! ---- ALE step drivers ------------------------------------------------
    INTEGER(KIND=C_INT32_T) FUNCTION bl_ale_active(h,ts) BIND(C,NAME="bl_ale_active")
      IMPORT :: C_INT32_T,C_PTR,bl_timestep_t
      TYPE(C_PTR),VALUE              :: h
      TYPE(bl_timestep_t),INTENT(IN) :: ts
    END FUNCTION bl_ale_active
    INTEGER(KIND=C_INT32_T) FUNCTION bl_ale_getmeshstatus(h)                    &
&                                          BIND(C,NAME="bl_ale_getmeshstatus")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: h
    END FUNCTION bl_ale_getmeshstatus
    INTEGER(KIND=C_INT32_T) FUNCTION bl_ale_getfluxvolume(h,ts)                 &
&                                          BIND(C,NAME="bl_ale_getfluxvolume")
      IMPORT :: C_INT32_T,C_PTR,bl_timestep_t
      TYPE(C_PTR),VALUE              :: h
      TYPE(bl_timestep_t),INTENT(IN) :: ts
    END FUNCTION bl_ale_getfluxvolume
    INTEGER(KIND=C_INT32_T) FUNCTION bl_ale_advect(h,ts,err) BIND(C,NAME="bl_ale_advect")
      IMPORT :: C_INT32_T,C_PTR,bl_timestep_t,bl_error_t
      TYPE(C_PTR),VALUE               :: h
      TYPE(bl_timestep_t),INTENT(IN)  :: ts
      TYPE(bl_error_t),   INTENT(OUT) :: err
    END FUNCTION bl_ale_advect
    INTEGER(KIND=C_INT32_T) FUNCTION bl_ale_hydroupdate(h,err)                  &
&                                          BIND(C,NAME="bl_ale_hydroupdate")
      IMPORT :: C_INT32_T,C_PTR,bl_error_t
      TYPE(C_PTR),VALUE             :: h
      TYPE(bl_error_t),INTENT(OUT)  :: err
    END FUNCTION bl_ale_hydroupdate

    ! ---- whole-step residency --------------------------------------------
    INTEGER(KIND=C_INT32_T) FUNCTION bl_run(h,nsteps,stop_at_time_end,ts,       &
&                                           dt_hist,idt_hist,ctl_hist,          &
&                                           steps_taken,err) BIND(C,NAME="bl_run")
      IMPORT :: C_INT32_T,C_PTR,C_DOUBLE,bl_timestep_t,bl_error_t
      TYPE(C_PTR),VALUE                                 :: h
      INTEGER(KIND=C_INT32_T),VALUE                     :: nsteps,stop_at_time_end
      TYPE(bl_timestep_t),                INTENT(INOUT) :: ts
      REAL(KIND=C_DOUBLE),    DIMENSION(*),INTENT(OUT)  :: dt_hist
      INTEGER(KIND=C_INT32_T),DIMENSION(*),INTENT(OUT)  :: idt_hist,ctl_hist
      INTEGER(KIND=C_INT32_T),            INTENT(OUT)   :: steps_taken
      TYPE(bl_error_t),                   INTENT(OUT)   :: err
    END FUNCTION bl_run
    INTEGER(KIND=C_INT32_T) FUNCTION bl_run_graph(h,nsteps,stop_at_time_end,ts, &
&                                                 dt_hist,idt_hist,ctl_hist,    &
&                                                 steps_taken,err)              &
&                                                 BIND(C,NAME="bl_run_graph")
      IMPORT :: C_INT32_T,C_PTR,C_DOUBLE,bl_timestep_t,bl_error_t
      TYPE(C_PTR),VALUE                                 :: h
      INTEGER(KIND=C_INT32_T),VALUE                     :: nsteps,stop_at_time_end
      TYPE(bl_timestep_t),                INTENT(INOUT) :: ts
      REAL(KIND=C_DOUBLE),    DIMENSION(*),INTENT(OUT)  :: dt_hist
      INTEGER(KIND=C_INT32_T),DIMENSION(*),INTENT(OUT)  :: idt_hist,ctl_hist
      INTEGER(KIND=C_INT32_T),            INTENT(OUT)   :: steps_taken
      TYPE(bl_error_t),                   INTENT(OUT)   :: err
    END FUNCTION bl_run_graph
    INTEGER(KIND=C_INT32_T) FUNCTION bl_totals(h,out) BIND(C,NAME="bl_totals")
      IMPORT :: C_INT32_T,C_PTR,bl_totals_t
      TYPE(C_PTR),VALUE              :: h
      TYPE(bl_totals_t),INTENT(OUT)  :: out
    END FUNCTION bl_totals
    ! region(nreg), material(nmat)
    INTEGER(KIND=C_INT32_T) FUNCTION bl_shortprint(h,region,material)           &
&                                    BIND(C,NAME="bl_shortprint")
      IMPORT :: C_INT32_T,C_PTR,bl_flag_totals_t
      TYPE(C_PTR),VALUE                               :: h
      TYPE(bl_flag_totals_t),DIMENSION(*),INTENT(OUT) :: region,material
    END FUNCTION bl_shortprint
    ! l1(1) density, l1(2) energy
    INTEGER(KIND=C_INT32_T) FUNCTION bl_sod_l1(h,time,x0,scale,l1)              &
&                                    BIND(C,NAME="bl_sod_l1")
      IMPORT :: C_INT32_T,C_PTR,C_DOUBLE
      TYPE(C_PTR),VALUE                       :: h
      REAL(KIND=C_DOUBLE),VALUE               :: time,x0,scale
      REAL(KIND=C_DOUBLE),DIMENSION(2),INTENT(OUT) :: l1
    END FUNCTION bl_sod_l1
    INTEGER(KIND=C_INT32_T) FUNCTION bl_set_diagnostics(h,on,x0,scale,maxrec)   &
&                                    BIND(C,NAME="bl_set_diagnostics")
      IMPORT :: C_INT32_T,C_PTR,C_DOUBLE
      TYPE(C_PTR),VALUE             :: h
      INTEGER(KIND=C_INT32_T),VALUE :: on,maxrec
      REAL(KIND=C_DOUBLE),VALUE     :: x0,scale
    END FUNCTION bl_set_diagnostics
    ! out(2+9*nreg+2, maxrec)
    INTEGER(KIND=C_INT32_T) FUNCTION bl_diagnostics(h,maxrec,out,nrec)          &
&                                    BIND(C,NAME="bl_diagnostics")
      IMPORT :: C_INT32_T,C_PTR,C_DOUBLE
      TYPE(C_PTR),VALUE                          :: h
      INTEGER(KIND=C_INT32_T),VALUE              :: maxrec
      REAL(KIND=C_DOUBLE),DIMENSION(*),INTENT(OUT) :: out
      INTEGER(KIND=C_INT32_T),INTENT(OUT)        :: nrec
    END FUNCTION bl_diagnostics

    ! ---- ensembles ---------------------------------------------------------
    INTEGER(KIND=C_INT32_T) FUNCTION bl_ensemble_create(inst,ninst,cfg,h)        &
&                                    BIND(C,NAME="bl_ensemble_create")
      IMPORT :: C_INT32_T,C_PTR,bl_instance_t,bl_config_t
      TYPE(bl_instance_t),DIMENSION(*),INTENT(IN) :: inst
      INTEGER(KIND=C_INT32_T),VALUE                :: ninst
      TYPE(bl_config_t),INTENT(IN)                 :: cfg
      TYPE(C_PTR),INTENT(OUT)                      :: h
    END FUNCTION bl_ensemble_create
    ! histories (ninst, nsteps): step-major as Fortran's (instance, step)
    INTEGER(KIND=C_INT32_T) FUNCTION bl_ensemble_run(h,nsteps,dt_hist,idt_hist,  &
&                                    ctl_hist,steps_taken,err)                   &
&                                    BIND(C,NAME="bl_ensemble_run")
      IMPORT :: C_INT32_T,C_PTR,C_DOUBLE,bl_error_t
      TYPE(C_PTR),VALUE                            :: h
      INTEGER(KIND=C_INT32_T),VALUE                :: nsteps
      REAL(KIND=C_DOUBLE),DIMENSION(*),INTENT(OUT)     :: dt_hist
      INTEGER(KIND=C_INT32_T),DIMENSION(*),INTENT(OUT) :: idt_hist,ctl_hist
      INTEGER(KIND=C_INT32_T),INTENT(OUT)          :: steps_taken
      TYPE(bl_error_t),INTENT(OUT)                 :: err
    END FUNCTION bl_ensemble_run
    INTEGER(KIND=C_INT32_T) FUNCTION bl_ensemble_offsets(h,eoff,noff)            &
&                                    BIND(C,NAME="bl_ensemble_offsets")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE                                :: h
      INTEGER(KIND=C_INT32_T),DIMENSION(*),INTENT(OUT) :: eoff,noff
    END FUNCTION bl_ensemble_offsets
    INTEGER(KIND=C_INT32_T) FUNCTION bl_ensemble_times(h,time,dt,nstep)          &
&                                    BIND(C,NAME="bl_ensemble_times")
      IMPORT :: C_INT32_T,C_PTR,C_DOUBLE
      TYPE(C_PTR),VALUE                                :: h
      REAL(KIND=C_DOUBLE),DIMENSION(*),INTENT(OUT)     :: time,dt
      INTEGER(KIND=C_INT32_T),DIMENSION(*),INTENT(OUT) :: nstep
    END FUNCTION bl_ensemble_times

    ! ---- node relaxation --------------------------------------------------
    INTEGER(KIND=C_INT32_T) FUNCTION bl_set_relaxation(h,rtype,nsweep,om,tol)   &
&                                    BIND(C,NAME="bl_set_relaxation")
      IMPORT :: C_INT32_T,C_PTR,C_DOUBLE
      TYPE(C_PTR),VALUE             :: h
      INTEGER(KIND=C_INT32_T),VALUE :: rtype,nsweep
      REAL(KIND=C_DOUBLE),VALUE     :: om,tol
    END FUNCTION bl_set_relaxation
    INTEGER(KIND=C_INT32_T) FUNCTION bl_relax_mesh(h,rtype,nsweep,om,tol)       &
&                                    BIND(C,NAME="bl_relax_mesh")
      IMPORT :: C_INT32_T,C_PTR,C_DOUBLE
      TYPE(C_PTR),VALUE             :: h
      INTEGER(KIND=C_INT32_T),VALUE :: rtype,nsweep
      REAL(KIND=C_DOUBLE),VALUE     :: om,tol
    END FUNCTION bl_relax_mesh

    ! ---- learned node relaxation (no counterpart in the reference) ---------
    INTEGER(KIND=C_INT32_T) FUNCTION bl_mlp_create(seed,capacity,m)             &
&                                    BIND(C,NAME="bl_mlp_create")
      IMPORT :: C_INT32_T,C_INT64_T,C_PTR
      INTEGER(KIND=C_INT64_T),VALUE :: seed
      INTEGER(KIND=C_INT32_T),VALUE :: capacity
      TYPE(C_PTR),INTENT(OUT)       :: m
    END FUNCTION bl_mlp_create
    INTEGER(KIND=C_INT32_T) FUNCTION bl_mlp_destroy(m) BIND(C,NAME="bl_mlp_destroy")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: m
    END FUNCTION bl_mlp_destroy
    INTEGER(KIND=C_INT32_T) FUNCTION bl_mlp_weights(m,wget,wset)                &
&                                    BIND(C,NAME="bl_mlp_weights")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: m,wget,wset
    END FUNCTION bl_mlp_weights
    INTEGER(KIND=C_INT32_T) FUNCTION bl_mlp_normalize(m) BIND(C,NAME="bl_mlp_normalize")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: m
    END FUNCTION bl_mlp_normalize
    INTEGER(KIND=C_INT32_T) FUNCTION bl_mlp_norm(m,nget,nset)                   &
&                                    BIND(C,NAME="bl_mlp_norm")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: m,nget,nset
    END FUNCTION bl_mlp_norm
    INTEGER(KIND=C_INT32_T) FUNCTION bl_mlp_forward(m,x,n,y)                    &
&                                    BIND(C,NAME="bl_mlp_forward")
      IMPORT :: C_INT32_T,C_PTR,C_FLOAT
      TYPE(C_PTR),VALUE                           :: m
      REAL(KIND=C_FLOAT),DIMENSION(*),INTENT(IN)  :: x
      INTEGER(KIND=C_INT32_T),VALUE               :: n
      REAL(KIND=C_FLOAT),DIMENSION(*),INTENT(OUT) :: y
    END FUNCTION bl_mlp_forward
    INTEGER(KIND=C_INT32_T) FUNCTION bl_mlp_add_pairs(m,x,t,n)                  &
&                                    BIND(C,NAME="bl_mlp_add_pairs")
      IMPORT :: C_INT32_T,C_PTR,C_FLOAT
      TYPE(C_PTR),VALUE                          :: m
      REAL(KIND=C_FLOAT),DIMENSION(*),INTENT(IN) :: x,t
      INTEGER(KIND=C_INT32_T),VALUE              :: n
    END FUNCTION bl_mlp_add_pairs
    INTEGER(KIND=C_INT32_T) FUNCTION bl_mlp_pairs(m,n,x,t)                      &
&                                    BIND(C,NAME="bl_mlp_pairs")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE                   :: m,x,t
      INTEGER(KIND=C_INT32_T),INTENT(OUT) :: n
    END FUNCTION bl_mlp_pairs
    INTEGER(KIND=C_INT32_T) FUNCTION bl_mlp_gradient(m,n,grad,loss)             &
&                                    BIND(C,NAME="bl_mlp_gradient")
      IMPORT :: C_INT32_T,C_PTR,C_FLOAT
      TYPE(C_PTR),VALUE                           :: m
      INTEGER(KIND=C_INT32_T),VALUE               :: n
      REAL(KIND=C_FLOAT),DIMENSION(*),INTENT(OUT) :: grad
      REAL(KIND=C_FLOAT),INTENT(OUT)              :: loss
    END FUNCTION bl_mlp_gradient
    INTEGER(KIND=C_INT32_T) FUNCTION bl_mlp_train(m,iters,batch,lr,seed,hist)   &
&                                    BIND(C,NAME="bl_mlp_train")
      IMPORT :: C_INT32_T,C_INT64_T,C_PTR,C_FLOAT
      TYPE(C_PTR),VALUE             :: m,hist
      INTEGER(KIND=C_INT32_T),VALUE :: iters,batch
      REAL(KIND=C_FLOAT),VALUE      :: lr
      INTEGER(KIND=C_INT64_T),VALUE :: seed
    END FUNCTION bl_mlp_train
    INTEGER(KIND=C_INT32_T) FUNCTION bl_relax_collect(h,m,stride)               &
&                                    BIND(C,NAME="bl_relax_collect")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE             :: h,m
      INTEGER(KIND=C_INT32_T),VALUE :: stride
    END FUNCTION bl_relax_collect
    INTEGER(KIND=C_INT32_T) FUNCTION bl_relax_learned(h,m)                      &
&                                    BIND(C,NAME="bl_relax_learned")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: h,m
    END FUNCTION bl_relax_learned
    INTEGER(KIND=C_INT32_T) FUNCTION bl_set_learned_check(h,max_aspect,relative) &
&                                    BIND(C,NAME="bl_set_learned_check")
      IMPORT :: C_INT32_T,C_PTR,C_DOUBLE
      TYPE(C_PTR),VALUE             :: h
      REAL(KIND=C_DOUBLE),VALUE     :: max_aspect
      INTEGER(KIND=C_INT32_T),VALUE :: relative
    END FUNCTION bl_set_learned_check
    INTEGER(KIND=C_INT32_T) FUNCTION bl_relax_learned_stats(h,steps,guarded,    &
&                                    fallbacks,reasons) BIND(C,NAME="bl_relax_learned_stats")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE                   :: h
      INTEGER(KIND=C_INT32_T),INTENT(OUT) :: steps,guarded,fallbacks,reasons
    END FUNCTION bl_relax_learned_stats
    INTEGER(KIND=C_INT32_T) FUNCTION bl_relax_features(h,feat)                  &
&                                    BIND(C,NAME="bl_relax_features")
      IMPORT :: C_INT32_T,C_PTR,C_FLOAT
      TYPE(C_PTR),VALUE                           :: h
      REAL(KIND=C_FLOAT),DIMENSION(*),INTENT(OUT) :: feat
    END FUNCTION bl_relax_features

    ! ---- kernel timing and diagnostics -----------------------------------
    INTEGER(KIND=C_INT32_T) FUNCTION bl_timing_enable(h,on) BIND(C,NAME="bl_timing_enable")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE             :: h
      INTEGER(KIND=C_INT32_T),VALUE :: on
    END FUNCTION bl_timing_enable
    INTEGER(KIND=C_INT32_T) FUNCTION bl_timing_reset(h) BIND(C,NAME="bl_timing_reset")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE :: h
    END FUNCTION bl_timing_reset
    INTEGER(KIND=C_INT32_T) FUNCTION bl_timing_query(h,name,avg_ms,launches)    &
&                                                    BIND(C,NAME="bl_timing_query")
      IMPORT :: C_INT32_T,C_INT64_T,C_PTR,C_DOUBLE,C_CHAR
      TYPE(C_PTR),VALUE                                 :: h
      CHARACTER(KIND=C_CHAR),DIMENSION(*),INTENT(IN)    :: name
      REAL(KIND=C_DOUBLE),                INTENT(OUT)   :: avg_ms
      INTEGER(KIND=C_INT64_T),            INTENT(OUT)   :: launches
    END FUNCTION bl_timing_query
    INTEGER(KIND=C_INT32_T) FUNCTION bl_timing_names(h,buf,buflen)              &
&                                                    BIND(C,NAME="bl_timing_names")
      IMPORT :: C_INT32_T,C_PTR,C_CHAR
      TYPE(C_PTR),VALUE                                 :: h
      CHARACTER(KIND=C_CHAR),DIMENSION(*),INTENT(OUT)   :: buf
      INTEGER(KIND=C_INT32_T),VALUE                     :: buflen
    END FUNCTION bl_timing_names
    INTEGER(KIND=C_INT32_T) FUNCTION bl_run_stats(h,moving,zeroed)              &
&                                                    BIND(C,NAME="bl_run_stats")
      IMPORT :: C_INT32_T,C_INT64_T,C_PTR
      TYPE(C_PTR),VALUE                                 :: h
      INTEGER(KIND=C_INT64_T),            INTENT(OUT)   :: moving,zeroed
    END FUNCTION bl_run_stats
    INTEGER(KIND=C_INT32_T) FUNCTION bl_autotune(h,reps,out)                    &
&                                                    BIND(C,NAME="bl_autotune")
      IMPORT :: C_INT32_T,C_PTR,bl_tuning_t
      TYPE(C_PTR),VALUE                                 :: h
      INTEGER(KIND=C_INT32_T),VALUE                     :: reps
      TYPE(bl_tuning_t),                  INTENT(OUT)   :: out
    END FUNCTION bl_autotune
    INTEGER(KIND=C_INT32_T) FUNCTION bl_set_tuning(h,t)                         &
&                                                    BIND(C,NAME="bl_set_tuning")
      IMPORT :: C_INT32_T,C_PTR,bl_tuning_t
      TYPE(C_PTR),VALUE                                 :: h
      TYPE(bl_tuning_t),                  INTENT(IN)    :: t
    END FUNCTION bl_set_tuning
    INTEGER(KIND=C_INT32_T) FUNCTION bl_get_tuning(h,t)                         &
&                                                    BIND(C,NAME="bl_get_tuning")
      IMPORT :: C_INT32_T,C_PTR,bl_tuning_t
      TYPE(C_PTR),VALUE                                 :: h
      TYPE(bl_tuning_t),                  INTENT(OUT)   :: t
    END FUNCTION bl_get_tuning
    INTEGER(KIND=C_INT32_T) FUNCTION bl_set_tiling(h,on) BIND(C,NAME="bl_set_tiling")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE             :: h
      INTEGER(KIND=C_INT32_T),VALUE :: on
    END FUNCTION bl_set_tiling
    INTEGER(KIND=C_INT32_T) FUNCTION bl_get_tiling(h,on,ntile,nboundary,nentries)&
&                                    BIND(C,NAME="bl_get_tiling")
      IMPORT :: C_INT32_T,C_PTR
      TYPE(C_PTR),VALUE                   :: h
      INTEGER(KIND=C_INT32_T),INTENT(OUT) :: on,ntile,nboundary,nentries
    END FUNCTION bl_get_tiling
    TYPE(C_PTR) FUNCTION bl_last_error_c() BIND(C,NAME="bl_last_error")
      IMPORT :: C_PTR
    END FUNCTION bl_last_error_c

  END INTERFACE

CONTAINS

  ! the library's last error message as a Fortran string
  FUNCTION bl_last_error() RESULT(msg)
    CHARACTER(LEN=256)                                :: msg
    TYPE(C_PTR)                                       :: p
    CHARACTER(KIND=C_CHAR),DIMENSION(:),POINTER       :: c
    INTEGER                                           :: i
    msg=' '
    p=bl_last_error_c()
    IF (.NOT.C_ASSOCIATED(p)) RETURN
    CALL C_F_POINTER(p,c,[LEN(msg)])
    DO i=1,LEN(msg)
      IF (c(i).EQ.C_NULL_CHAR) EXIT
      msg(i:i)=c(i)
    ENDDO
  END FUNCTION bl_last_error

  ! error_t%serr text of a bl_error_t
  FUNCTION bl_error_text(err) RESULT(msg)
    TYPE(bl_error_t),INTENT(IN) :: err
    CHARACTER(LEN=BL_SLEN)      :: msg
    INTEGER                     :: i
    DO i=1,BL_SLEN
      msg(i:i)=err%serr(i)
    ENDDO
  END FUNCTION bl_error_text

END MODULE bookleaf_hip_mod
