! fortran_host.f90 -- a Fortran host driving libbookleaf_hip.so through
! bookleaf_hip_mod, the way a BookLeaf build with the drop-in drivers would.
!
! Reads a problem in the stream layout of oracle/refio.py write_problem,
! uploads it into the library's dataAPI mirror, initialises (connectivity,
! geometry, init_dr_state, hydro_dr_init), then runs
!   - nhost steps through the per-driver entry points in the order of
!     inf_solver_getdt / inf_solver_lagstep / inf_solver_alestep
!     (inf_solver_getdt.f90:35-70, inf_solver_lagstep.f90:44-129,
!     inf_solver_alestep.f90:39-74) with time_dr_calc/time_dr_end done here
!     (time_dr_advance.f90:38-132), and
!   - the remaining steps through bl_run_graph,
! and writes dt, idtel and the dt control of every step followed by the
! final eldensity, elenergy, elpressure, ndx, ndy, ndu, ndv.
!
! usage: fortran_host <problem.bin> <out.bin> <nhost>
PROGRAM fortran_host

  USE, INTRINSIC :: iso_c_binding, ONLY: C_INT32_T,C_INT64_T,C_DOUBLE,C_PTR,    &
&                                        C_LOC
  USE bookleaf_hip_mod

  IMPLICIT NONE

  INTEGER,PARAMETER :: ELDENSITY=1,ELENERGY=2,ELPRESSURE=3,NDX=35,NDY=36,      &
&                      NDU=37,NDV=38,INDTYPE=39,IELND=41,IELMAT=44,IELREG=45
  CHARACTER(LEN=512)                           :: fin,fout,arg
  INTEGER(KIND=C_INT32_T)                      :: hdr(11),nel,nnd,nmat,nreg,   &
&                                                 nsteps,nhost,rc,taken,k
  REAL(KIND=C_DOUBLE)                          :: rcfg(16)
  REAL(KIND=C_DOUBLE),   ALLOCATABLE,TARGET    :: kap(:),pme(:),epar(:),       &
&                                                 x(:),y(:),u(:),v(:),rho(:),  &
&                                                 en(:),pre(:),dth(:)
  INTEGER(KIND=C_INT32_T),ALLOCATABLE,TARGET   :: zdt(:),zml(:),ityp(:),       &
&                                                 ielnd4(:),ireg(:),imat(:),   &
&                                                 ityp_nd(:),idth(:),ctlh(:)
  TYPE(bl_sizes_t)                             :: sz
  TYPE(bl_config_t)                            :: cfg
  TYPE(bl_timestep_t)                          :: ts
  TYPE(bl_error_t)                             :: err
  TYPE(C_PTR)                                  :: h
  TYPE(bl_flag_totals_t),ALLOCATABLE           :: regtab(:),mattab(:)

  CALL GET_COMMAND_ARGUMENT(1,fin)
  CALL GET_COMMAND_ARGUMENT(2,fout)
  CALL GET_COMMAND_ARGUMENT(3,arg)
  READ(arg,*) nhost

  ! ---- problem file --------------------------------------------------------
  OPEN(UNIT=10,FILE=TRIM(fin),ACCESS='STREAM',FORM='UNFORMATTED',STATUS='OLD')
  READ(10) hdr
  nel=hdr(1); nnd=hdr(2); nmat=hdr(3); nreg=hdr(4); nsteps=hdr(5)
  READ(10) rcfg
  ALLOCATE(kap(nreg),pme(nreg),zdt(nreg),zml(nreg),ityp(0:nmat),              &
&          epar(6*(nmat+1)))
  READ(10) kap,pme,zdt,zml
  ityp(0)=0
  epar(:)=0.0_C_DOUBLE
  READ(10) ityp(1:nmat)
  READ(10) epar(7:6*(nmat+1))
  ALLOCATE(ielnd4(4*nel),ireg(nel),imat(nel),ityp_nd(nnd),x(nnd),y(nnd),      &
&          u(nnd),v(nnd),rho(nel),en(nel),pre(nel))
  READ(10) ielnd4,ireg,imat,ityp_nd,x,y,u,v,rho,en
  CLOSE(10)

  ! ---- sizes and configuration (the rationalised values of the deck) ------
  sz=bl_sizes_t(nel,nel,nel,nnd,nnd,nnd,MAX(nel,nnd),nmat,nreg,0,0,0,0)
  cfg%zcut=rcfg(1);  cfg%zerocut=rcfg(2); cfg%dencut=rcfg(3); cfg%accut=rcfg(4)
  cfg%cvisc1=rcfg(5); cfg%cvisc2=rcfg(6); cfg%cfl_sf=rcfg(7); cfg%div_sf=rcfg(8)
  cfg%ccut=rcfg(9);  cfg%pcut=rcfg(10); cfg%time_start=rcfg(11)
  cfg%time_end=rcfg(12); cfg%dt_g=rcfg(13); cfg%dt_min=rcfg(14)
  cfg%dt_max=rcfg(15); cfg%dt_initial=rcfg(16)
  cfg%zhg=hdr(10); cfg%zsp=hdr(11); cfg%ztq=0; cfg%pad0=0
  cfg%zeul=hdr(8); cfg%adv_type=hdr(9); cfg%zon=0; cfg%sf=0.5_C_DOUBLE
  cfg%zexist=cfg%zeul
  cfg%maxtime=HUGE(1.0_C_DOUBLE)
  IF (cfg%zeul.NE.0) THEN
    cfg%mintime=-HUGE(1.0_C_DOUBLE)
  ELSE
    cfg%mintime=HUGE(1.0_C_DOUBLE)
  ENDIF

  CALL chk(bl_create(sz,h),'bl_create')
  CALL chk(bl_set_config(h,cfg),'bl_set_config')
  CALL chk(bl_set_regions(h,nreg,kap,pme,zdt,zml),'bl_set_regions')
  CALL chk(bl_set_eos(h,nmat,ityp,epar),'bl_set_eos')
  CALL chk(bl_upload(h,IELND,C_LOC(ielnd4(1)),INT(4*nel,C_INT64_T)),'ielnd')
  CALL chk(bl_upload(h,IELREG,C_LOC(ireg(1)),INT(nel,C_INT64_T)),'ielreg')
  CALL chk(bl_upload(h,IELMAT,C_LOC(imat(1)),INT(nel,C_INT64_T)),'ielmat')
  CALL chk(bl_upload(h,INDTYPE,C_LOC(ityp_nd(1)),INT(nnd,C_INT64_T)),'indtype')
  CALL chk(bl_upload(h,NDX,C_LOC(x(1)),INT(nnd,C_INT64_T)),'ndx')
  CALL chk(bl_upload(h,NDY,C_LOC(y(1)),INT(nnd,C_INT64_T)),'ndy')
  CALL chk(bl_upload(h,NDU,C_LOC(u(1)),INT(nnd,C_INT64_T)),'ndu')
  CALL chk(bl_upload(h,NDV,C_LOC(v(1)),INT(nnd,C_INT64_T)),'ndv')
  CALL chk(bl_upload(h,ELDENSITY,C_LOC(rho(1)),INT(nel,C_INT64_T)),'density')
  CALL chk(bl_upload(h,ELENERGY,C_LOC(en(1)),INT(nel,C_INT64_T)),'energy')

  ! ---- inf_init: connectivity, geometry, state, hydro ----------------------
  CALL chk(bl_init_mesh(h),'bl_init_mesh')
  CALL chk(bl_geometry_getgeometry(h,err),'bl_geometry_getgeometry')
  CALL chk(bl_init_state(h),'bl_init_state')
  CALL chk(bl_hydro_init(h),'bl_hydro_init')

  ! the timestep before step 1: dt = dts = dt_initial, control INITIAL
  ts=bl_timestep_t(cfg%time_start,cfg%dt_initial,cfg%dt_initial,0,0,0,         &
&                  BL_DT_INITIAL)
  ALLOCATE(dth(nsteps),idth(nsteps),ctlh(nsteps))
  dth(:)=0.0_C_DOUBLE; idth(:)=0; ctlh(:)=0

  ! ---- host-sequenced steps ------------------------------------------------
  DO k=1,MIN(nhost,nsteps)
    CALL getdt()
    dth(k)=ts%dt; idth(k)=ts%idtel; ctlh(k)=ts%sdt
    CALL lagstep()
    IF (bl_ale_active(h,ts).NE.0) CALL alestep()
  ENDDO

  ! ---- the rest on the device-resident runner -----------------------------
  IF (nsteps.GT.nhost) THEN
    rc=bl_run_graph(h,nsteps-nhost,0,ts,dth(nhost+1:),idth(nhost+1:),          &
&                   ctlh(nhost+1:),taken,err)
    CALL chk(rc,'bl_run_graph')
  ENDIF

  CALL chk(bl_download(h,ELDENSITY,C_LOC(rho(1)),INT(nel,C_INT64_T)),'rho')
  CALL chk(bl_download(h,ELENERGY,C_LOC(en(1)),INT(nel,C_INT64_T)),'en')
  CALL chk(bl_download(h,ELPRESSURE,C_LOC(pre(1)),INT(nel,C_INT64_T)),'pre')
  CALL chk(bl_download(h,NDX,C_LOC(x(1)),INT(nnd,C_INT64_T)),'x')
  CALL chk(bl_download(h,NDY,C_LOC(y(1)),INT(nnd,C_INT64_T)),'y')
  CALL chk(bl_download(h,NDU,C_LOC(u(1)),INT(nnd,C_INT64_T)),'u')
  CALL chk(bl_download(h,NDV,C_LOC(v(1)),INT(nnd,C_INT64_T)),'v')
  ! hydro_dr_shortprint's tables
  ALLOCATE(regtab(nreg),mattab(nmat))
  CALL chk(bl_shortprint(h,regtab,mattab),'bl_shortprint')
  OPEN(UNIT=11,FILE=TRIM(fout),ACCESS='STREAM',FORM='UNFORMATTED',            &
&      STATUS='REPLACE')
  WRITE(11) nsteps,dth,idth,ctlh,rho,en,pre,x,y,u,v
  WRITE(11) regtab,mattab
  CLOSE(11)
  CALL chk(bl_destroy(h),'bl_destroy')
  WRITE(*,'(A,I0,A)') 'fortran_host: ',nsteps,' steps done'

CONTAINS

  SUBROUTINE chk(rc,what)
    INTEGER(KIND=C_INT32_T),INTENT(IN) :: rc
    CHARACTER(LEN=*),       INTENT(IN) :: what
    IF (rc.NE.BL_SUCCESS) THEN
      WRITE(*,'(A,A,A,A)') 'fortran_host: ',what,' failed: ',TRIM(bl_last_error())
      STOP 1
    ENDIF
  END SUBROUTINE chk

  ! inf_solver_getdt + time_dr_calc/time_dr_end: strict '<' minimum over
  ! MAXIMUM, GROWTH, INITIAL (first step), CFL, DIV in that order
  SUBROUTINE getdt()
    TYPE(bl_dt_t)       :: cfl,div
    REAL(KIND=C_DOUBLE) :: best,c
    INTEGER             :: bidx,bsdt
    CALL chk(bl_hydro_getdt(h,cfl,div,err),'bl_hydro_getdt')
    best=cfg%dt_max; bidx=-1; bsdt=BL_DT_MAXIMUM
    c=cfg%dt_g*ts%dt
    IF (c.LT.best) THEN
      best=c; bidx=-1; bsdt=BL_DT_GROWTH
    ENDIF
    IF (ts%nstep.EQ.0.AND.cfg%dt_initial.LT.best) THEN
      best=cfg%dt_initial; bidx=-1; bsdt=BL_DT_INITIAL
    ENDIF
    IF (cfl%rdt.LT.best) THEN
      best=cfl%rdt; bidx=cfl%idt; bsdt=cfl%sdt
    ENDIF
    IF (div%rdt.LT.best) THEN
      best=div%rdt; bidx=div%idt; bsdt=div%sdt
    ENDIF
    ts%nstep=ts%nstep+1
    ts%dt=best; ts%idtel=bidx; ts%sdt=bsdt
    ts%time=ts%time+best
    IF (best.LT.cfg%dt_min) THEN
      WRITE(*,'(A)') 'fortran_host: dt < dt_min'
      STOP 1
    ENDIF
  END SUBROUTINE getdt

  ! inf_solver_lagstep: predictor, then corrector
  SUBROUTINE lagstep()
    ts%zcorrector=0; ts%dts=0.5_C_DOUBLE*ts%dt
    CALL chk(bl_hydro_setpredictor(h),'setpredictor')
    CALL chk(bl_hydro_getforce(h,ts),'getforce')
    CALL chk(bl_hydro_getstate(h,ts,err),'getstate')
    CALL chk(bl_hydro_getenergy(h,ts),'getenergy')
    CALL chk(bl_hydro_geteos(h),'geteos')
    ts%zcorrector=1; ts%dts=ts%dt
    CALL chk(bl_hydro_setcorrector(h),'setcorrector')
    CALL chk(bl_hydro_getforce(h,ts),'getforce')
    CALL chk(bl_hydro_initacceleration(h),'initacceleration')
    CALL chk(bl_hydro_scatteracceleration(h),'scatteracceleration')
    CALL chk(bl_hydro_getacceleration(h),'getacceleration')
    CALL chk(bl_hydro_applyacceleration(h,ts),'applyacceleration')
    CALL chk(bl_hydro_getstate(h,ts,err),'getstate')
    CALL chk(bl_hydro_getenergy(h,ts),'getenergy')
    CALL chk(bl_hydro_geteos(h),'geteos')
  END SUBROUTINE lagstep

  ! inf_solver_alestep
  SUBROUTINE alestep()
    CALL chk(bl_ale_getmeshstatus(h),'ale_getmeshstatus')
    CALL chk(bl_ale_getfluxvolume(h,ts),'ale_getfluxvolume')
    CALL chk(bl_ale_advect(h,ts,err),'ale_advect')
    CALL chk(bl_ale_hydroupdate(h,err),'ale_hydroupdate')
  END SUBROUTINE alestep

END PROGRAM fortran_host
