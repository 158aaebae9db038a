! ref_driver.f90 -- host program for the reference build in oracle/_ref.
!
! TEST INFRASTRUCTURE ONLY.  Links the reference's own solver, driver and
! kernel objects (compiled from /root/reference/src by oracle/build_ref.py)
! and replaces only the deck reader / mesh generator / IC setup with a binary
! problem file, so that the reference and the HIP implementation start from
! identical state.  The set-up calls mirror inf_init (inf_init.F90:53-101):
! defaults, rationalise, data_set, connectivity, node types, ielsort,
! geometry, init_dr_state, hydro_dr_init.  The step loop is the reference's
! inf_control_solver loop (inf_control_solver.f90:53-69) calling the
! reference's inf_solver_getdt / inf_solver_lagstep / inf_solver_alestep.
!
! Built with -DBL_DROPIN (oracle/build_ref.py --dropin), the reference's
! hot-path driver modules are replaced by bookleaf_f90_amd/fortran/dropin/
! (same module and procedure names), so the same inf_solver_* loop runs the
! step on the GPU; every dump first brings the device fields back
! (bookleaf_dropin_sync).
!
! usage: bookleaf_ref <problem.bin> <dump.bin> [mode]
!   mode 0 (default): full steps, dump every "dump_every" steps
!   mode 1: granular -- step 1 only, dump after every driver call
!   mode 2: timing -- no dumps, print element-steps/s of the step loop
PROGRAM bl_ref_driver

  USE, INTRINSIC :: iso_c_binding, ONLY: C_F_POINTER,C_LOC
  USE dataAPI_kinds_mod,       ONLY: ink,rlk,lok,lak
  USE dataAPI_params_mod,      ONLY: SUCCESS,NCORN
  USE dataAPI_types_mod,       ONLY: config_t,runtime_t,data_t,error_t,dt_t
  USE dataAPI_id_mod
  USE timerAPI_types_mod,      ONLY: timer_t
  USE timer_control_mod,       ONLY: timer_init
  USE inf_init_mod,            ONLY: inf_init_parallel,inf_init_default
  USE init_dr_parallel_mod,    ONLY: init_dr_mpi_comm
  USE setup_read_mod,          ONLY: setup_timestep_rationalise,               &
&                                    setup_utils_rationalise
  USE time_cf_mod,             ONLY: time_cf_rationalise
  USE hydro_cf_mod,            ONLY: hydro_cf_rationalise
  USE ale_cf_mod,              ONLY: ale_cf_rationalise
  USE eos_cf_mod,              ONLY: eos_cf_rationalise
  USE data_mod,                ONLY: data_set,data_reset_cpquant,             &
&                                    data_reset_mxquant
  USE init_kn_mod,             ONLY: init_kn_serialghosts,                     &
&                                    init_kn_connectivity,init_kn_nodetype
  USE init_dr_utils_mod,       ONLY: init_dr_utils,init_dr_state
  USE hydro_dr_init_mod,       ONLY: hydro_dr_init
  USE geometry_dr_volume_mod,  ONLY: geometry_dr_getgeometry
  USE inf_solver_getdt_mod,    ONLY: inf_solver_getdt
  USE inf_solver_lagstep_mod,  ONLY: inf_solver_lagstep
  USE inf_solver_alestep_mod,  ONLY: inf_solver_alestep
  USE time_dr_utils_mod,       ONLY: time_finish
  USE time_dr_advance_mod,     ONLY: time_dr_calc,time_dr_end
  USE time_dr_set_mod,         ONLY: time_dr_setpredictor,time_dr_setcorrector
  USE hydro_dr_getdt_mod,      ONLY: hydro_dr_getdt
  USE hydro_dr_set_mod,        ONLY: hydro_dr_setpredictor,                    &
&                                    hydro_dr_setcorrector
  USE hydro_dr_getforce_mod,   ONLY: hydro_dr_getforce
  USE hydro_dr_getstate_mod,   ONLY: hydro_dr_getstate
  USE hydro_dr_getenergy_mod,  ONLY: hydro_dr_getenergy
  USE hydro_dr_geteos_mod,     ONLY: hydro_dr_geteos
  USE hydro_dr_getacceleration_mod,ONLY: hydro_dr_initacceleration,            &
&                                    hydro_dr_scatteracceleration,             &
&                                    hydro_dr_getacceleration,                 &
&                                    hydro_dr_applyacceleration
  USE ale_dr_utils_mod,        ONLY: ale_active
  USE ale_dr_getmeshstatus_mod,ONLY: ale_dr_getmeshstatus
  USE ale_dr_getfluxvolume_mod,ONLY: ale_dr_getfluxvolume
  USE ale_dr_advect_mod,       ONLY: ale_dr_advect
  USE ale_dr_update_mod,       ONLY: ale_dr_hydroupdate
  USE timerAPI_id_mod
  USE hydro_kn_print_mod,      ONLY: hydro_kn_init_shortprint,                 &
&                                    hydro_kn_calc_shortprint
  USE utils_kn_gather_mod,     ONLY: utils_kn_cngather,utils_kn_mxgathercncp
#ifdef BL_DROPIN
  USE bookleaf_dropin_mod,     ONLY: bookleaf_dropin_sync,bookleaf_dropin_detach
#endif

  IMPLICIT NONE

  CHARACTER(LEN=512)                       :: fin,fout,smode,sdeck
  TYPE(config_t), POINTER                  :: config
  TYPE(runtime_t)                          :: runtime
  TYPE(data_t),   DIMENSION(:),ALLOCATABLE :: dh
  TYPE(timer_t),  DIMENSION(:),ALLOCATABLE :: timer
  TYPE(error_t)                            :: error
  INTEGER(KIND=ink) :: nel,nnd,nmat,nreg,nsteps,dump_every,stop_end,zeul,     &
&                      adv_type,mode,istep,ii,ndump,iu,io
  INTEGER(KIND=ink) :: zhgflag,zspflag,ncp,nmx,mcp,mmx
  INTEGER(KIND=ink),DIMENSION(:),  ALLOCATABLE :: imxel,imxfcp,imxncp,icpmat,  &
&                                                 icpnext,icpprev
  REAL(KIND=rlk),   DIMENSION(:),  ALLOCATABLE :: frvol,cpden,cpene
  REAL(KIND=rlk)    :: cut(4),hyd(4),ecut(2),tim(6)
  REAL(KIND=rlk),   DIMENSION(:),  ALLOCATABLE :: kappareg,pmeritreg
  INTEGER(KIND=ink),DIMENSION(:),  ALLOCATABLE :: zdtnotreg,zmidlength,itype
  REAL(KIND=rlk),   DIMENSION(:,:),ALLOCATABLE :: param
  INTEGER(KIND=ink),DIMENSION(:),  ALLOCATABLE :: ielnd,ielreg,ielmat,indtype
  REAL(KIND=rlk),   DIMENSION(:),  ALLOCATABLE :: ndx,ndy,ndu,ndv,den,ene
  INTEGER(KIND=8)   :: c0,c1,crate
  REAL(KIND=rlk)    :: wall
  REAL(KIND=rlk),   DIMENSION(:),POINTER :: rp
  INTEGER(KIND=ink),DIMENSION(:),POINTER :: ip

  CALL GET_COMMAND_ARGUMENT(1,fin)
  CALL GET_COMMAND_ARGUMENT(2,fout)
  mode=0_ink
  IF (COMMAND_ARGUMENT_COUNT().GE.3) THEN
    CALL GET_COMMAND_ARGUMENT(3,smode)
    READ(smode,*) mode
  ENDIF

  ! ---- read the problem file (layout written by tests/refio.py) ----------
  OPEN(NEWUNIT=iu,FILE=TRIM(fin),ACCESS='STREAM',FORM='UNFORMATTED',          &
&      STATUS='OLD',ACTION='READ')
  READ(iu) nel,nnd,nmat,nreg,nsteps,dump_every,stop_end,zeul,adv_type,         &
&          zhgflag,zspflag
  READ(iu) cut,hyd,ecut,tim
  ALLOCATE(kappareg(nreg),pmeritreg(nreg),zdtnotreg(nreg),zmidlength(nreg))
  READ(iu) kappareg,pmeritreg,zdtnotreg,zmidlength
  ALLOCATE(itype(nmat),param(6,nmat))
  READ(iu) itype,param
  ALLOCATE(ielnd(NCORN*nel),ielreg(nel),ielmat(nel),indtype(nnd))
  ALLOCATE(ndx(nnd),ndy(nnd),ndu(nnd),ndv(nnd),den(nel),ene(nel))
  READ(iu) ielnd,ielreg,ielmat,indtype,ndx,ndy,ndu,ndv,den,ene
  ! mixed cells: the flattened lists of mix_dr_flatten (mix_dr_list.f90)
  READ(iu,IOSTAT=io) ncp,nmx,mcp,mmx
  IF (io.NE.0) THEN
    ncp=0_ink
    nmx=0_ink
  ENDIF
  IF (ncp.GT.0_ink) THEN
    ALLOCATE(imxel(nmx),imxfcp(nmx),imxncp(nmx),icpmat(ncp),icpnext(ncp),      &
&            icpprev(ncp),frvol(ncp),cpden(ncp),cpene(ncp))
    READ(iu) imxel,imxfcp,imxncp,icpmat,icpnext,icpprev,frvol,cpden,cpene
  ENDIF
  CLOSE(iu)

  ! ---- configuration: main + inf_init_default + inf_init_rationalise -----
  ALLOCATE(config)
  CALL inf_init_parallel(config%comm)
  CALL timer_init(timer,error)
  CALL inf_init_default(sdeck,config,runtime,timer,dh)
  CALL init_dr_mpi_comm(config%comm)

  config%global%zcut=cut(1)
  config%global%zerocut=cut(2)
  config%global%dencut=cut(3)
  config%global%accut=cut(4)
  config%hydro%cvisc1=hyd(1)
  config%hydro%cvisc2=hyd(2)
  config%hydro%cfl_sf=hyd(3)
  config%hydro%div_sf=hyd(4)
  config%eos%ccut=ecut(1)
  config%eos%pcut=ecut(2)
  config%time%time_start=tim(1)
  config%time%time_end=tim(2)
  config%time%dt_g=tim(3)
  config%time%dt_min=tim(4)
  config%time%dt_max=tim(5)
  config%time%dt_initial=tim(6)
  config%ale%zeul=(zeul.NE.0_ink)
  config%ale%adv_type=adv_type
  DEALLOCATE(config%hydro%kappareg,config%hydro%pmeritreg,                     &
&            config%hydro%zdtnotreg,config%hydro%zmidlength)
  ALLOCATE(config%hydro%kappareg(nreg),config%hydro%pmeritreg(nreg),           &
&          config%hydro%zdtnotreg(nreg),config%hydro%zmidlength(nreg))
  config%hydro%kappareg(:)=kappareg(:)
  config%hydro%pmeritreg(:)=pmeritreg(:)
  config%hydro%zdtnotreg(:)=(zdtnotreg(:).NE.0_ink)
  config%hydro%zmidlength(:)=(zmidlength(:).NE.0_ink)
  DEALLOCATE(config%eos%itype,config%eos%param)
  ALLOCATE(config%eos%itype(nmat),config%eos%param(6,nmat))
  config%eos%itype(:)=itype(:)
  config%eos%param(:,:)=param(:,:)
  DEALLOCATE(config%io%sregion,config%io%smaterial)
  ALLOCATE(config%io%sregion(nreg),config%io%smaterial(nmat))
  config%io%sregion(:)='region'
  config%io%smaterial(:)='material'

  runtime%sizes%nel=nel
  runtime%sizes%nnd=nnd
  runtime%sizes%nmat=nmat
  runtime%sizes%nreg=nreg
  ! void material 0 (setup_read.f90:458, eos_config.f90:163-226)
  CALL eos_cf_rationalise(config%eos,runtime%sizes,error)
  CALL check(error)
  CALL init_kn_serialghosts(runtime%sizes%nel,runtime%sizes%nel1,              &
&                           runtime%sizes%nel2,runtime%sizes%nnd,              &
&                           runtime%sizes%nnd1,runtime%sizes%nnd2)
  CALL setup_timestep_rationalise(config%time,runtime%timestep)
  CALL setup_utils_rationalise(config,runtime%sizes,error)
  CALL check(error)
  CALL time_cf_rationalise(config%time,config%comm%spatial,config%io,error)
  CALL check(error)
  CALL hydro_cf_rationalise(config%hydro,config%comm%spatial,config%eos,       &
&                           config%io,config%global,runtime%sizes,error)
  CALL check(error)
  CALL ale_cf_rationalise(config%ale,config%time,config%global,               &
&                         config%comm%spatial,config%eos,error)
  CALL check(error)

  ! ---- data registry and mesh: init_dr_mesh (init_dr_mesh.F90:41-97) ----
  CALL data_set(config,runtime%sizes,dh,error)
  CALL check(error)
  ! data_entry registers shape-[1] pointers with the true length in dsize
  ! (data.f90:428-453); view them at full length
  CALL iview(ielndid,ip)
  ip(:)=ielnd(:)
  CALL iview(ielregid,ip)
  ip(:)=ielreg(:)
  CALL iview(ielmatid,ip)
  ip(:)=ielmat(:)
  CALL iview(indtypeid,ip)
  ip(:)=indtype(:)
  CALL rview(ndxid,rp)
  rp(:)=ndx(:)
  CALL rview(ndyid,rp)
  rp(:)=ndy(:)
  CALL init_kn_connectivity(runtime%sizes%nel2,dh(ielndid)%iaddr,              &
&                           dh(ielelid)%iaddr,dh(ielfcid)%iaddr)
  CALL init_kn_nodetype(runtime%sizes%nel1,runtime%sizes%nnd1,                 &
&                       dh(ielndid)%iaddr,dh(indtypeid)%iaddr)
  CALL init_dr_utils(config,runtime%sizes,dh,error)
  CALL check(error)

  ! ---- mixed cells: what material_shape + mix_dr_flatten leave
  ! (setup_IC.f90:1585-1634, mix_dr_list.f90) -------------------------------
  IF (ncp.GT.0_ink) THEN
    ! capacities as mix_dr_addel / mix_dr_addcp size them (mix_dr_list.f90)
    CALL data_reset_mxquant(mmx,runtime%sizes,dh,io)
    CALL data_reset_cpquant(mcp,runtime%sizes,dh,io)
    runtime%sizes%nmx=nmx
    runtime%sizes%ncp=ncp
    CALL iview(imxelid,ip)
    ip(1:nmx)=imxel(:)
    CALL iview(imxfcpid,ip)
    ip(1:nmx)=imxfcp(:)
    CALL iview(imxncpid,ip)
    ip(1:nmx)=imxncp(:)
    CALL iview(icpmatid,ip)
    ip(1:ncp)=icpmat(:)
    CALL iview(icpnextid,ip)
    ip(1:ncp)=icpnext(:)
    CALL iview(icpprevid,ip)
    ip(1:ncp)=icpprev(:)
    CALL rview(frvolumeid,rp)
    rp(1:ncp)=frvol(:)
    CALL rview(cpdensityid,rp)
    rp(1:ncp)=cpden(:)
    CALL rview(cpenergyid,rp)
    rp(1:ncp)=cpene(:)
  ENDIF

  ! ---- initial conditions (the part setup_IC would do) + init_dr_state ---
  CALL rview(nduid,rp)
  rp(:)=ndu(:)
  CALL rview(ndvid,rp)
  rp(:)=ndv(:)
  CALL rview(eldensityid,rp)
  rp(:)=den(:)
  CALL rview(elenergyid,rp)
  rp(:)=ene(:)
  CALL geometry_dr_getgeometry(runtime%sizes,timer(TGETGEOMETRYIID),dh,error)
  CALL check(error)
  CALL init_dr_state(config,runtime%sizes,timer(TGETEOSIID),dh)
  CALL hydro_dr_init(config%hydro,runtime%sizes,dh)

  ! ---- step loop ---------------------------------------------------------
  OPEN(NEWUNIT=iu,FILE=TRIM(fout),ACCESS='STREAM',FORM='UNFORMATTED',         &
&      STATUS='REPLACE',ACTION='WRITE')
  ndump=0
  IF (mode.EQ.0_ink) CALL dump(0_ink)

  IF (mode.EQ.1_ink) THEN
    CALL granular_step()
  ELSE
    CALL SYSTEM_CLOCK(c0,crate)
    DO istep=1,nsteps
      CALL inf_solver_getdt(config,runtime,timer,dh)
      CALL inf_solver_lagstep(config,runtime,timer,dh)
      CALL inf_solver_alestep(config,runtime,timer,dh)
      IF (mode.EQ.0_ink) THEN
        WRITE(iu) -1_ink,runtime%timestep%nstep,runtime%timestep%idtel,        &
&                 runtime%timestep%dt,runtime%timestep%time,                   &
&                 runtime%timestep%sdt
        IF (dump_every.GT.0_ink) THEN
          IF (MOD(istep,dump_every).EQ.0_ink) CALL dump(istep)
        ENDIF
      ENDIF
      IF (stop_end.NE.0_ink) THEN
        IF (time_finish(runtime%timestep,config%time)) EXIT
      ENDIF
    ENDDO
    CALL SYSTEM_CLOCK(c1)
    wall=REAL(c1-c0,rlk)/REAL(crate,rlk)
    IF (mode.EQ.0_ink) THEN
      CALL dump(runtime%timestep%nstep)
      CALL shortprint_record()
    ENDIF
    IF (mode.EQ.2_ink) THEN
      WRITE(*,'(a,i12,a,i8,a,es16.9,a,es16.9)') 'REFTIMING nel=',nel,         &
&       ' steps=',runtime%timestep%nstep,' seconds=',wall,                     &
&       ' elsteps_per_s=',REAL(nel,rlk)*REAL(runtime%timestep%nstep,rlk)/wall
    ENDIF
  ENDIF
  WRITE(iu) -9_ink
  CLOSE(iu)
#ifdef BL_DROPIN
  CALL bookleaf_dropin_detach()
#endif

CONTAINS

  SUBROUTINE check(err)
    TYPE(error_t),INTENT(IN) :: err
    IF (err%ierr.NE.SUCCESS) THEN
      WRITE(*,*) 'REFERROR ',err%ierr,TRIM(err%serr)
      STOP 3
    ENDIF
  END SUBROUTINE check

  ! One dump record: marker 1, label, then every allocated dataAPI field.
  SUBROUTINE dump(label)
    INTEGER(KIND=ink),INTENT(IN) :: label
    INTEGER(KIND=ink) :: id,knd
    LOGICAL(KIND=lak),DIMENSION(:),POINTER :: zp
#ifdef BL_DROPIN
    ! the drop-in drivers keep the fields on the GPU: bring them back first
    CALL bookleaf_dropin_sync(dh)
#endif
    WRITE(iu) 1_ink,label,runtime%timestep%nstep,runtime%timestep%idtel,       &
&             runtime%timestep%dt,runtime%timestep%time,runtime%timestep%dts
    DO id=1,SIZE(dh)
      IF (ASSOCIATED(dh(id)%raddr)) THEN
        knd=1
        CALL rview(id,rp)
        WRITE(iu) id,knd,INT(SIZE(rp),KIND=8)
        WRITE(iu) rp
      ELSEIF (ASSOCIATED(dh(id)%iaddr)) THEN
        knd=2
        CALL iview(id,ip)
        WRITE(iu) id,knd,INT(SIZE(ip),KIND=8)
        WRITE(iu) ip
      ELSEIF (ASSOCIATED(dh(id)%zaddr)) THEN
        knd=3
        CALL C_F_POINTER(C_LOC(dh(id)%zaddr(1)),zp,[dh(id)%dsize])
        WRITE(iu) id,knd,INT(SIZE(zp),KIND=8)
        WRITE(iu) zp
      ENDIF
    ENDDO
    WRITE(iu) 0_ink,0_ink,0_8
    ndump=ndump+1
  END SUBROUTINE dump

  ! The tables of hydro_dr_shortprint (before its averaging and printing),
  ! from the reference's own kernels on local arrays: marker 2, nreg, nmat,
  ! then for the regions and then the materials the nine columns volume,
  ! mass, ie, ke, pressure, pmin, pmax, dmin, dmax, each over the flags.
  SUBROUTINE shortprint_record()
    INTEGER(KIND=ink) :: nr,nm,ne,nn,nc,nx
    REAL(KIND=rlk),DIMENSION(runtime%sizes%nreg) :: rv,rm,rk,ri,rdn,rdx,rp,    &
&                                                   rpn,rpx
    REAL(KIND=rlk),DIMENSION(runtime%sizes%nmat) :: mv,mm,mk,mi,mdn,mdx,mp,    &
&                                                   mpn,mpx
    REAL(KIND=rlk),DIMENSION(:,:),ALLOCATABLE :: cu,cv,pw,pu,pv
    nr=runtime%sizes%nreg
    nm=runtime%sizes%nmat
    ne=runtime%sizes%nel
    nn=runtime%sizes%nnd
    nc=runtime%sizes%ncp
    nx=runtime%sizes%nmx
    ALLOCATE(cu(NCORN,ne),cv(NCORN,ne))
    CALL utils_kn_cngather(ne,nn,dh(ielndid)%iaddr,dh(nduid)%raddr,cu)
    CALL utils_kn_cngather(ne,nn,dh(ielndid)%iaddr,dh(ndvid)%raddr,cv)
    CALL hydro_kn_init_shortprint(nr,rv,rm,rk,ri,rdn,rdx,rp,rpn,rpx)
    CALL hydro_kn_calc_shortprint(nr,ne,config%hydro%global%dencut,            &
&        dh(ielregid)%iaddr,dh(eldensityid)%raddr,dh(elenergyid)%raddr,        &
&        dh(elpressureid)%raddr,dh(elvolumeid)%raddr,dh(elmassid)%raddr,       &
&        dh(cnwtid)%raddr,cu,cv,rv,rm,rk,ri,rdn,rdx,rp,rpn,rpx)
    CALL hydro_kn_init_shortprint(nm,mv,mm,mk,mi,mdn,mdx,mp,mpn,mpx)
    CALL hydro_kn_calc_shortprint(nm,ne,config%hydro%global%dencut,            &
&        dh(ielmatid)%iaddr,dh(eldensityid)%raddr,dh(elenergyid)%raddr,        &
&        dh(elpressureid)%raddr,dh(elvolumeid)%raddr,dh(elmassid)%raddr,       &
&        dh(cnwtid)%raddr,cu,cv,mv,mm,mk,mi,mdn,mdx,mp,mpn,mpx)
    IF (nc.GT.0_ink) THEN
      ALLOCATE(pw(NCORN,nc),pu(NCORN,nc),pv(NCORN,nc))
      CALL utils_kn_mxgathercncp(ne,nx,nc,dh(imxelid)%iaddr,dh(imxfcpid)%iaddr,&
&          dh(imxncpid)%iaddr,dh(cnwtid)%raddr,pw)
      CALL utils_kn_mxgathercncp(ne,nx,nc,dh(imxelid)%iaddr,dh(imxfcpid)%iaddr,&
&          dh(imxncpid)%iaddr,cu,pu)
      CALL utils_kn_mxgathercncp(ne,nx,nc,dh(imxelid)%iaddr,dh(imxfcpid)%iaddr,&
&          dh(imxncpid)%iaddr,cv,pv)
      CALL hydro_kn_calc_shortprint(nm,nc,config%hydro%global%dencut,          &
&          dh(icpmatid)%iaddr,dh(cpdensityid)%raddr,dh(cpenergyid)%raddr,      &
&          dh(cppressureid)%raddr,dh(cpvolumeid)%raddr,dh(cpmassid)%raddr,     &
&          pw,pu,pv,mv,mm,mk,mi,mdn,mdx,mp,mpn,mpx)
      DEALLOCATE(pw,pu,pv)
    ENDIF
    DEALLOCATE(cu,cv)
    WRITE(iu) 2_ink,nr,nm
    WRITE(iu) rv,rm,ri,rk,rp,rpn,rpx,rdn,rdx
    WRITE(iu) mv,mm,mi,mk,mp,mpn,mpx,mdn,mdx
  END SUBROUTINE shortprint_record

  SUBROUTINE rview(id,p)
    INTEGER(KIND=ink),                     INTENT(IN)  :: id
    REAL(KIND=rlk),DIMENSION(:),POINTER,   INTENT(OUT) :: p
    CALL C_F_POINTER(C_LOC(dh(id)%raddr(1)),p,[dh(id)%dsize])
  END SUBROUTINE rview

  SUBROUTINE iview(id,p)
    INTEGER(KIND=ink),                     INTENT(IN)  :: id
    INTEGER(KIND=ink),DIMENSION(:),POINTER,INTENT(OUT) :: p
    CALL C_F_POINTER(C_LOC(dh(id)%iaddr(1)),p,[dh(id)%dsize])
  END SUBROUTINE iview

  ! One step, calling the drivers in the order of inf_solver_getdt,
  ! inf_solver_lagstep and inf_solver_alestep, with a dump after each call.
  ! Labels: 100 getdt, 200..207 predictor, 300..311 corrector, 400.. ALE.
  SUBROUTINE granular_step()
    TYPE(dt_t),POINTER :: first,current
    CALL dump(0_ink)
    NULLIFY(first,current)
    CALL time_dr_calc(config%time,runtime%timestep,timer(TGETDTID),first,      &
&                     current)
    CALL hydro_dr_getdt(config%hydro,runtime%sizes,timer,dh,current,error)
    CALL check(error)
    CALL time_dr_end(config%time,runtime,timer,dh,first,current,error)
    CALL check(error)
    CALL dump(100_ink)
    ! predictor
    CALL time_dr_setpredictor(runtime%timestep)
    CALL hydro_dr_setpredictor(runtime%sizes,dh)
    CALL dump(200_ink)
    CALL hydro_dr_getforce(config%hydro,runtime,timer,dh)
    CALL dump(201_ink)
    CALL hydro_dr_getstate(runtime,timer(TGETGEOMETRYLID),dh,error)
    CALL check(error)
    CALL dump(202_ink)
    CALL hydro_dr_getenergy(config%hydro,runtime,dh)
    CALL dump(203_ink)
    CALL hydro_dr_geteos(config%hydro,runtime%sizes,timer,dh)
    CALL dump(204_ink)
    ! corrector
    CALL time_dr_setcorrector(runtime%timestep)
    CALL hydro_dr_setcorrector(runtime%sizes,dh)
    CALL dump(300_ink)
    CALL hydro_dr_getforce(config%hydro,runtime,timer,dh)
    CALL dump(301_ink)
    CALL hydro_dr_initacceleration(config%hydro%comm,runtime%sizes,           &
&                                  timer(TCOMMLID),dh)
    CALL hydro_dr_scatteracceleration(config%hydro,runtime%sizes,dh)
    CALL dump(302_ink)
    CALL hydro_dr_getacceleration(config%hydro,runtime%sizes,dh)
    CALL dump(303_ink)
    CALL hydro_dr_applyacceleration(runtime,dh)
    CALL dump(304_ink)
    CALL hydro_dr_getstate(runtime,timer(TGETGEOMETRYLID),dh,error)
    CALL check(error)
    CALL dump(305_ink)
    CALL hydro_dr_getenergy(config%hydro,runtime,dh)
    CALL dump(306_ink)
    CALL hydro_dr_geteos(config%hydro,runtime%sizes,timer,dh)
    CALL dump(307_ink)
    ! ALE
    IF (ale_active(config%ale,runtime%timestep)) THEN
      CALL ale_dr_getmeshstatus(config%ale,runtime%sizes,                      &
&                               timer(TALEGETMESHSTATUSID),dh)
      CALL dump(400_ink)
      CALL ale_dr_getfluxvolume(config%ale,runtime,timer(TALEGETFLUXVOLUMEID), &
&                               dh)
      CALL dump(401_ink)
      CALL ale_dr_advect(config%ale,runtime,timer,dh,error)
      CALL check(error)
      CALL dump(402_ink)
      CALL ale_dr_hydroupdate(config%ale,runtime%sizes,timer,dh,error)
      CALL check(error)
      CALL dump(403_ink)
    ENDIF
  END SUBROUTINE granular_step

END PROGRAM bl_ref_driver
