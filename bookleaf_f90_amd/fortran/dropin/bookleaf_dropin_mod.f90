! bookleaf_dropin_mod.f90 -- the library handle behind BookLeaf's dh(:)
! registry, shared by the drop-in driver modules of this directory.
!
! The other files here replace the reference's hot-path driver modules
! (hydro_dr_getdt_mod ... ale_dr_update_mod: same module names, procedure
! names and argument lists as src/packages/hydro/driver/*.f90 and
! src/packages/ale/driver/*.f90), so that the reference's own solver
! (inf_solver_getdt / lagstep / alestep and the inf_solver_* they call)
! links against them unchanged and runs the step on the GPU through
! libbookleaf_hip.so.
!
! Residency.  The first call of any of the replaced drivers creates the
! handle for the run's sizes_t, copies the configuration (global_t,
! hydro_t, eos_t, and ale_t when an ALE driver passes it) and the region and
! material tables, uploads every allocated dh(id) array and builds the
! node->corner maps from ielnd/ielsort1/ielsort2.  From then on the device
! holds the fields: the drivers move no data across PCIe.  A host that reads
! dh between steps (output, printing, checks) calls bookleaf_dropin_sync
! first; a host that writes a field after the first driver call passes it
! back with bookleaf_dropin_push.  bookleaf_dropin_detach frees the handle.
!
! Errors are reported through error_t as the reference's drivers report
! them (ierr, iout, serr).  A failure of the library itself (no device,
! out of memory) in a driver that has no error_t stops the program with the
! library's message, as an allocation failure in the reference would.
MODULE bookleaf_dropin_mod

  USE, INTRINSIC :: iso_c_binding, ONLY: C_PTR,C_NULL_PTR,C_ASSOCIATED,C_LOC,  &
&                                        C_INT32_T,C_INT64_T,C_NULL_CHAR
  USE dataAPI_kinds_mod, ONLY: ink,rlk,lok
  USE dataAPI_params_mod,ONLY: SUCCESS,FAILURE,SLEN
  USE dataAPI_types_mod, ONLY: data_t,sizes_t,error_t,timestep_t,hydro_t,ale_t
  USE bookleaf_hip_mod,  ONLY: bl_sizes_t,bl_config_t,bl_timestep_t,bl_error_t,&
&                              bl_create,bl_destroy,bl_set_config,             &
&                              bl_set_regions,bl_set_eos,bl_field_info,        &
&                              bl_upload,bl_download,bl_build_maps,            &
&                              bl_last_error,BL_SUCCESS,BL_DT_MAXIMUM,         &
&                              BL_DT_GROWTH,BL_DT_INITIAL,BL_DT_CFL,BL_DT_DIV

  IMPLICIT NONE

  PRIVATE

  TYPE(C_PTR),      SAVE :: handle=C_NULL_PTR
  TYPE(bl_config_t),SAVE :: cfg
  LOGICAL,          SAVE :: ale_known=.FALSE.

  PUBLIC :: dropin_attach,dropin_ale,dropin_handle,dropin_timestep,           &
&           dropin_error,dropin_require,bookleaf_dropin_sync,                  &
&           bookleaf_dropin_push,bookleaf_dropin_detach

CONTAINS

  ! The handle of this process (attached by the first driver call).
  FUNCTION dropin_handle() RESULT(h)
    TYPE(C_PTR) :: h
    IF (.NOT.C_ASSOCIATED(handle)) CALL fail('a BookLeaf driver ran before '   &
&     //'hydro_dr_getdt attached the GPU handle')
    h=handle
  END FUNCTION dropin_handle

  ! Create the handle and move the host registry to the device, once.
  SUBROUTINE dropin_attach(hydro,sizes,dh)
    TYPE(hydro_t),             INTENT(IN) :: hydro
    TYPE(sizes_t),             INTENT(IN) :: sizes
    TYPE(data_t), DIMENSION(:),INTENT(IN) :: dh
    TYPE(bl_sizes_t) :: bs
    INTEGER(KIND=ink) :: nreg,nmat,ir
    INTEGER(KIND=C_INT32_T),DIMENSION(:),ALLOCATABLE :: zdt,zmid

    IF (C_ASSOCIATED(handle)) RETURN
    bs%nel=sizes%nel
    bs%nel1=sizes%nel1
    bs%nel2=sizes%nel2
    bs%nnd=sizes%nnd
    bs%nnd1=sizes%nnd1
    bs%nnd2=sizes%nnd2
    bs%nsz=sizes%nsz
    bs%nmat=sizes%nmat
    bs%nreg=sizes%nreg
    bs%ncp=sizes%ncp
    bs%nmx=sizes%nmx
    bs%mcp=sizes%mcp
    bs%mmx=sizes%mmx
    CALL check(bl_create(bs,handle),'bl_create')

    ! global_t, hydro_t and eos_t; time_t is not needed by the drivers, ale_t
    ! comes with the first ALE driver call
    cfg%zcut=hydro%global%zcut
    cfg%zerocut=hydro%global%zerocut
    cfg%dencut=hydro%global%dencut
    cfg%accut=hydro%global%accut
    cfg%cvisc1=hydro%cvisc1
    cfg%cvisc2=hydro%cvisc2
    cfg%cfl_sf=hydro%cfl_sf
    cfg%div_sf=hydro%div_sf
    cfg%zhg=flag(hydro%zhg)
    cfg%zsp=flag(hydro%zsp)
    cfg%ztq=flag(hydro%ztq)
    cfg%pad0=0
    cfg%ccut=hydro%eos%ccut
    cfg%pcut=hydro%eos%pcut
    cfg%time_start=0.0_rlk
    cfg%time_end=HUGE(1.0_rlk)
    cfg%dt_g=1.0_rlk
    cfg%dt_min=0.0_rlk
    cfg%dt_max=HUGE(1.0_rlk)
    cfg%dt_initial=HUGE(1.0_rlk)
    cfg%mintime=HUGE(1.0_rlk)
    cfg%maxtime=HUGE(1.0_rlk)
    cfg%sf=0.0_rlk
    cfg%adv_type=1
    cfg%zexist=0
    cfg%zon=0
    cfg%zeul=0
    CALL check(bl_set_config(handle,cfg),'bl_set_config')

    nreg=sizes%nreg
    ALLOCATE(zdt(nreg),zmid(nreg))
    DO ir=1,nreg
      zdt(ir)=flag(hydro%zdtnotreg(ir))
      zmid(ir)=flag(hydro%zmidlength(ir))
    ENDDO
    CALL check(bl_set_regions(handle,nreg,hydro%kappareg,hydro%pmeritreg,zdt,  &
&                             zmid),'bl_set_regions')
    DEALLOCATE(zdt,zmid)
    ! itype(0:nmat), param(:,0:nmat) after eos_cf_rationalise
    nmat=sizes%nmat
    CALL check(bl_set_eos(handle,nmat,hydro%eos%itype,hydro%eos%param),       &
&              'bl_set_eos')

    CALL push_all(dh)
    CALL check(bl_build_maps(handle),'bl_build_maps')
  END SUBROUTINE dropin_attach

  ! ale_t's scalars, from the ALE drivers' own argument.
  SUBROUTINE dropin_ale(ale)
    TYPE(ale_t),INTENT(IN) :: ale
    INTEGER(KIND=C_INT32_T) :: zex,zon,zeu
    zex=flag(ale%zexist)
    zon=flag(ale%zon)
    zeu=flag(ale%zeul)
    IF (ale_known.AND.cfg%adv_type.EQ.ale%adv_type.AND.cfg%zexist.EQ.zex.AND.  &
&       cfg%zon.EQ.zon.AND.cfg%zeul.EQ.zeu.AND.cfg%mintime.EQ.ale%mintime.AND. &
&       cfg%maxtime.EQ.ale%maxtime.AND.cfg%sf.EQ.ale%sf) RETURN
    cfg%adv_type=ale%adv_type
    cfg%zexist=zex
    cfg%zon=zon
    cfg%zeul=zeu
    cfg%mintime=ale%mintime
    cfg%maxtime=ale%maxtime
    cfg%sf=ale%sf
    CALL check(bl_set_config(dropin_handle(),cfg),'bl_set_config')
    ale_known=.TRUE.
  END SUBROUTINE dropin_ale

  ! timestep_t (dataAPI_types.f90:46-52) as the library's bl_timestep_t
  FUNCTION dropin_timestep(ts) RESULT(bt)
    TYPE(timestep_t),INTENT(IN) :: ts
    TYPE(bl_timestep_t) :: bt
    bt%time=ts%time
    bt%dt=ts%dt
    bt%dts=ts%dts
    bt%nstep=ts%nstep
    bt%idtel=ts%idtel
    bt%zcorrector=flag(ts%zcorrector)
    SELECT CASE(TRIM(ADJUSTL(ts%sdt)))
      CASE('MAXIMUM')
        bt%sdt=BL_DT_MAXIMUM
      CASE('GROWTH')
        bt%sdt=BL_DT_GROWTH
      CASE('INITIAL')
        bt%sdt=BL_DT_INITIAL
      CASE('CFL')
        bt%sdt=BL_DT_CFL
      CASE('DIV')
        bt%sdt=BL_DT_DIV
      CASE DEFAULT
        bt%sdt=0
    END SELECT
  END FUNCTION dropin_timestep

  ! the library's bl_error_t into error_t: SUCCESS when rc is BL_SUCCESS,
  ! else the code, halt level and message the library filled in (or its
  ! last error message when it filled in none)
  SUBROUTINE dropin_error(rc,berr,error)
    INTEGER(KIND=C_INT32_T),INTENT(IN)  :: rc
    TYPE(bl_error_t),       INTENT(IN)  :: berr
    TYPE(error_t),          INTENT(OUT) :: error
    INTEGER :: ii
    error%ierr=SUCCESS
    error%iout=0_ink
    error%serr=' '
    IF (rc.EQ.BL_SUCCESS) RETURN
    error%ierr=berr%ierr
    error%iout=berr%iout
    DO ii=1,MIN(SLEN,SIZE(berr%serr))
      IF (berr%serr(ii).EQ.C_NULL_CHAR) EXIT
      error%serr(ii:ii)=berr%serr(ii)
    ENDDO
    IF (error%ierr.EQ.SUCCESS) THEN
      error%ierr=FAILURE
      error%serr=last_error()
    ENDIF
  END SUBROUTINE dropin_error

  ! a library call of a driver without error_t
  SUBROUTINE dropin_require(rc,what)
    INTEGER(KIND=C_INT32_T),INTENT(IN) :: rc
    CHARACTER(LEN=*),       INTENT(IN) :: what
    CALL check(rc,what)
  END SUBROUTINE dropin_require

  ! Every allocated dh(id) array from the device into host memory.
  SUBROUTINE bookleaf_dropin_sync(dh)
    TYPE(data_t),DIMENSION(:),INTENT(INOUT) :: dh
    INTEGER(KIND=ink) :: id
    IF (.NOT.C_ASSOCIATED(handle)) RETURN
    DO id=1,SIZE(dh)
      IF (ASSOCIATED(dh(id)%raddr)) THEN
        CALL check(bl_download(handle,id,C_LOC(dh(id)%raddr(1)),                &
&                              count_of(dh(id),id)),'bl_download')
      ELSEIF (ASSOCIATED(dh(id)%iaddr)) THEN
        CALL check(bl_download(handle,id,C_LOC(dh(id)%iaddr(1)),                &
&                              count_of(dh(id),id)),'bl_download')
      ELSEIF (ASSOCIATED(dh(id)%zaddr)) THEN
        CALL check(bl_download(handle,id,C_LOC(dh(id)%zaddr(1)),                &
&                              count_of(dh(id),id)),'bl_download')
      ENDIF
    ENDDO
  END SUBROUTINE bookleaf_dropin_sync

  ! One host-written dh(id) array back to the device (no-op before the
  ! handle exists: attaching uploads everything).
  SUBROUTINE bookleaf_dropin_push(dh,id)
    TYPE(data_t),DIMENSION(:),INTENT(IN) :: dh
    INTEGER(KIND=ink),        INTENT(IN) :: id
    IF (.NOT.C_ASSOCIATED(handle)) RETURN
    CALL push_one(dh,id)
  END SUBROUTINE bookleaf_dropin_push

  SUBROUTINE bookleaf_dropin_detach()
    IF (.NOT.C_ASSOCIATED(handle)) RETURN
    CALL check(bl_destroy(handle),'bl_destroy')
    handle=C_NULL_PTR
    ale_known=.FALSE.
  END SUBROUTINE bookleaf_dropin_detach

  ! ---- internals ----------------------------------------------------------

  SUBROUTINE push_all(dh)
    TYPE(data_t),DIMENSION(:),INTENT(IN) :: dh
    INTEGER(KIND=ink) :: id
    DO id=1,SIZE(dh)
      CALL push_one(dh,id)
    ENDDO
  END SUBROUTINE push_all

  SUBROUTINE push_one(dh,id)
    TYPE(data_t),DIMENSION(:),INTENT(IN) :: dh
    INTEGER(KIND=ink),        INTENT(IN) :: id
    IF (ASSOCIATED(dh(id)%raddr)) THEN
      CALL check(bl_upload(handle,id,C_LOC(dh(id)%raddr(1)),count_of(dh(id),id)),&
&                'bl_upload')
    ELSEIF (ASSOCIATED(dh(id)%iaddr)) THEN
      CALL check(bl_upload(handle,id,C_LOC(dh(id)%iaddr(1)),count_of(dh(id),id)),&
&                'bl_upload')
    ELSEIF (ASSOCIATED(dh(id)%zaddr)) THEN
      CALL check(bl_upload(handle,id,C_LOC(dh(id)%zaddr(1)),count_of(dh(id),id)),&
&                'bl_upload')
    ENDIF
  END SUBROUTINE push_one

  ! the entries to move: the host's registered length (data_entry's dsize),
  ! which must fit the device mirror of that id
  FUNCTION count_of(d,id) RESULT(n)
    TYPE(data_t),     INTENT(IN) :: d
    INTEGER(KIND=ink),INTENT(IN) :: id
    INTEGER(KIND=C_INT64_T) :: n,dev
    INTEGER(KIND=C_INT32_T) :: knd
    n=INT(d%dsize,KIND=C_INT64_T)
    CALL check(bl_field_info(handle,id,knd,dev),'bl_field_info')
    IF (n.GT.dev) CALL fail('field '//TRIM(d%dname)//' is longer on the host '  &
&                           //'than its device mirror')
  END FUNCTION count_of

  INTEGER(KIND=C_INT32_T) FUNCTION flag(z)
    LOGICAL(KIND=lok),INTENT(IN) :: z
    flag=0
    IF (z) flag=1
  END FUNCTION flag

  SUBROUTINE check(rc,what)
    INTEGER(KIND=C_INT32_T),INTENT(IN) :: rc
    CHARACTER(LEN=*),       INTENT(IN) :: what
    IF (rc.NE.BL_SUCCESS) CALL fail(what//': '//last_error())
  END SUBROUTINE check

  SUBROUTINE fail(msg)
    CHARACTER(LEN=*),INTENT(IN) :: msg
    WRITE(*,'(a)') 'ERROR: bookleaf_dropin: '//msg
    ERROR STOP 1
  END SUBROUTINE fail

  FUNCTION last_error() RESULT(s)
    CHARACTER(LEN=SLEN) :: s
    s=bl_last_error()
  END FUNCTION last_error

END MODULE bookleaf_dropin_mod
