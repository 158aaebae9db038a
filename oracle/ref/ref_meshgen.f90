! ref_meshgen -- host program around the reference's mesh generator.
!
! TEST INFRASTRUCTURE ONLY.  Linked by oracle/build_ref.py with the
! reference's own objects (compiled from /root/reference/src) into
! oracle/_ref/bookleaf_meshgen.  It reads the &MESH namelist of a deck with
! setup_mesh_gen (setup_mesh.F90:75-104: input, checks, and the region
! meshes with their SOR relaxation, :593-1089), transfers the mesh with
! setup_mesh_transfer (:1091-1250) and writes it out, so that the oracle's
! relaxation stencil can be checked against the generator's converged
! meshes.
!
! usage: bookleaf_meshgen <deck> <mesh.bin>
! output (stream, native endian): nel, nnd (int32); ielnd(4,nel),
! indtype(nnd), ielreg(nel) (int32); ndx(nnd), ndy(nnd) (real64)
PROGRAM bl_ref_meshgen

  USE dataAPI_kinds_mod,  ONLY: ink,rlk
  USE dataAPI_params_mod, ONLY: SUCCESS,NCORN
  USE dataAPI_types_mod,  ONLY: global_t,sizes_t,error_t
  USE timerAPI_types_mod, ONLY: timer_t
  USE global_cf_mod,      ONLY: global_cf_defaults
  USE setup_mesh_mod,     ONLY: setup_mesh_gen,setup_mesh_transfer

  IMPLICIT NONE

  CHARACTER(LEN=512)                           :: deck,fout
  TYPE(global_t)                               :: global
  TYPE(sizes_t)                                :: sizes
  TYPE(timer_t)                                :: timer
  TYPE(error_t)                                :: error
  INTEGER(KIND=ink)                            :: iu
  INTEGER(KIND=ink),DIMENSION(:),  ALLOCATABLE :: indtype,ielreg,ielmat
  INTEGER(KIND=ink),DIMENSION(:,:),ALLOCATABLE :: ielnd
  REAL(KIND=rlk),   DIMENSION(:),  ALLOCATABLE :: ndx,ndy

  CALL GET_COMMAND_ARGUMENT(1,deck)
  CALL GET_COMMAND_ARGUMENT(2,fout)

  CALL global_cf_defaults(global)
  sizes%nmat=1_ink
  timer%zactive=.FALSE.
  timer%time=0.0_rlk

  CALL setup_mesh_gen(TRIM(deck),global,sizes,timer,error)
  CALL check()
  ALLOCATE(indtype(sizes%nnd),ndx(sizes%nnd),ndy(sizes%nnd))
  ALLOCATE(ielreg(sizes%nel),ielmat(sizes%nel),ielnd(NCORN,sizes%nel))
  CALL setup_mesh_transfer(timer,sizes%nel,sizes%nnd,indtype,ndx,ndy,ielreg,   &
&                          ielmat,ielnd,error)
  CALL check()

  OPEN(NEWUNIT=iu,FILE=TRIM(fout),ACCESS='STREAM',FORM='UNFORMATTED',         &
&      STATUS='REPLACE',ACTION='WRITE')
  WRITE(iu) sizes%nel,sizes%nnd
  WRITE(iu) ielnd,indtype,ielreg
  WRITE(iu) ndx,ndy
  CLOSE(iu)

CONTAINS

  SUBROUTINE check()
    IF (error%ierr.NE.SUCCESS) THEN
      WRITE(*,*) 'REFERROR ',error%ierr,TRIM(error%serr)
      STOP 3
    ENDIF
  END SUBROUTINE check

END PROGRAM bl_ref_meshgen
