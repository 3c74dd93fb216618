! Replacements for the bodies of the reference's OPDPTH (src/lblrtm.f90:
! 6135) and XMERGE (src/xmerge.f90:25), compiled with the reference's
! double-precision flags (-fdefault-integer-8 -fdefault-real-8
! -fdefault-double-8) and linked ahead of its objects (the originals
! weakened), so that its driver XLAYER calls liblblx through lblx_c.
!
! They read the reference's COMMON blocks (declared here with the
! reference's layouts, which a drop-in must share) and keep its file
! contract: OPDPTH writes the layer's optical depth to KFILE as HIRAC1 and
! PANEL do (file header, then per panel a header record and the panel's
! values, src/oprop.f90:405-414, 1524-1526); XMERGE writes the merged
! radiance and transmittance to MFILE as EMINIT/RADMRG and EMOUT do
! (src/xmerge.f90:2401, 3737-3790).  Between the two the layer's optical
! depth and the merged result of the layers below stay in memory (the
! latter on the device), so XMERGE does not read KFILE or LFILE back.
!
! Supported: IHIRAC = 1 with no line-file reuse (ILNFLG = 0), IOD = 0,
! no cross sections; XMERGE with IEMIT = 1, IPATHL = 1 or 3 (upwelling or downwelling), no
! aerosols, a constant surface emissivity and no surface-property file.
! Anything else stops with the reason, as the reference stops on errors.

module lblx_shim_state
  use iso_c_binding
  use lblx_c
  implicit none
  type(c_ptr), save :: ctx = c_null_ptr
  ! the last OPDPTH result: its layer and run, its optical depth
  type(lblx_run), save :: od_run
  type(lblx_layer), save :: od_layer
  real(c_double), allocatable, target, save :: od(:)
  logical, save :: od_valid = .false.
  ! the merged result of the layers below (device) and the layer it is on
  type(lblx_layer), target, save :: old_layer
  type(c_ptr), save :: d_rold = c_null_ptr, d_told = c_null_ptr
  logical, save :: old_valid = .false.
  ! 1 where the executable is built from src/oprop_voigt.f90 (the tabulated
  ! exact-Voigt build), set by the caller before the first OPDPTH
  integer(c_int32_t), save :: voigt_build = 0
contains
  subroutine lblx_stop(where, rc)
    character(len=*), intent(in) :: where
    integer(c_int), intent(in) :: rc
    write (*, '(a,a,a,i0,a)') ' ', where, ': liblblx error ', rc, ' ('//lblx_message(rc)//')'
    stop 'liblblx error'
  end subroutine lblx_stop

  subroutine lblx_context()
    integer(c_int) :: rc
    if (c_associated(ctx)) return
    rc = lblx_create(ctx, 0_c_int)
    if (rc /= LBLX_OK) call lblx_stop('lblx_create', rc)
    ! LINFIL is TAPE3 in the working directory (src/lblrtm.f90:599-617)
    rc = lblx_load_tape3(ctx, 'TAPE3'//c_null_char)
    if (rc /= LBLX_OK) call lblx_stop('lblx_load_tape3', rc)
  end subroutine lblx_context
end module lblx_shim_state

! One sequential unformatted record of nwords default words, the record
! BUFOUT writes.
subroutine lblx_bufout(iunit, words, nwords)
  implicit none
  integer :: iunit, nwords
  real :: words(nwords)
  write (iunit) words
end subroutine lblx_bufout

SUBROUTINE OPDPTH (MPTS)
  use lblx_shim_state
  IMPLICIT REAL*8 (V)
  COMMON /MANE/ P0,TEMP0,NLAYER,DVXM,H2OSLF,WTOT,ALBAR,ADBAR,AVBAR,    &
       &        AVFIX,LAYRFX,SECNT0,SAMPLE,DVSET,ALFAL0,AVMASS,          &
       &        DPTMIN,DPTFAC,ALTAV,AVTRAT,TDIFF1,TDIFF2,ALTD1,          &
       &        ALTD2,ANGLE,IANT,LTGNT,LH1,LH2,IPFLAG,PLAY,TLAY,         &
       &        EXTID(10)
  CHARACTER*8 EXTID
  character*8 XID, HMOLID, YID
  real*8 SECANT, XALTZ
  COMMON /FILHDR/ XID(10),SECANT,PAVE,TAVE,HMOLID(60),XALTZ(4),        &
       &          WK(60),PZL,PZU,TZL,TZU,WBROAD,DV ,V1 ,V2 ,TBOUND,      &
       &          EMISIV,FSCDID(17),NMOL,LAYRS ,YI1,YID(10),LSTWDF
  COMMON /IFIL/ IRD,IPR,IPU,NOPR,NFHDRF,NPHDRF,NFHDRL,NPHDRL,          &
       &        NLNGTH,KFILE,KPANEL,LINFIL,NFILE,IAFIL,IEXFIL,           &
       &        NLTEFL,LNFIL4,LNGTH4,IBRD
  COMMON /IODFLG/ DVOUT
  COMMON /RCNTRL/ ILNFLG
  COMMON /CNTSCL/ XSELF,XFRGN,XCO2C,XO3CN,XO2CN,XN2CN,XRAYL
  COMMON /CLOSURE/ FRGNX
  character*1 FRGNX
  DIMENSION FILHDR(2)
  EQUIVALENCE (FSCDID(1),IHIRAC) , (FSCDID(2),ILBLF4),                 &
       &      (FSCDID(3),IXSCNT) , (FSCDID(4),IAERSL),                   &
       &      (FSCDID(5),IEMIT) , (FSCDID(8),IPATHL),                    &
       &      (FSCDID(9),JRAD) , (XID(1),FILHDR(1))
  type(lblx_run) :: r
  integer(c_int) :: rc
  integer(c_int64_t) :: n
  real(c_double) :: v1p0
  real(c_double), allocatable :: v1p(:), v2p(:)
  integer(c_int32_t), allocatable :: nlim(:)
  real(c_double) :: q1(1)
  integer(c_int32_t) :: qn(1)
  integer(c_int) :: np
  integer :: k, i, o, m
  REAL*8 V1PK, V2PK
  REAL DVP
  INTEGER NLIMK

  if (IHIRAC /= 1 .or. ILNFLG /= 0 .or. DVOUT /= 0. .or. IXSCNT / 10 /= 0) &
       stop 'OPDPTH: option outside the liblblx path (IHIRAC, ILNFLG, IOD, IXSECT)'
  call lblx_context()
  ! record 1.2/1.3 as OPDPTH sees them; the ICNTNM codes are already the
  ! /CNTSCL/ factors the driver set (src/lblrtm.f90:688-760)
  r%v1 = V1
  r%v2 = V2
  r%sample = SAMPLE
  r%alfal0 = ALFAL0
  r%avmass = AVMASS
  r%dptmin = DPTMIN
  IF (IEMIT.EQ.0.AND.IPATHL.EQ.2) r%dptmin = 2.*DPTMIN
  r%dptfac = DPTFAC
  r%ilblf4 = int(ILBLF4, c_int32_t)
  r%icntnm = 0
  IF (MOD(IXSCNT, 10).NE.0) r%icntnm = 6
  r%irad = 0
  r%precision = LBLX_FP64
  r%cntscl = [XSELF, XFRGN, XCO2C, XO3CN, XO2CN, XN2CN, XRAYL]
  r%ifrgnx = int(ichar(FRGNX) - ichar('0'), c_int32_t)
  r%iaersl = int(IAERSL, c_int32_t)
  r%voigt = voigt_build
  r%ibrd = int(IBRD, c_int32_t)
  od_layer%pave = PAVE
  od_layer%tave = TAVE
  od_layer%wk = 0.
  m = min(NMOL, LBLX_MXMOL)
  od_layer%wk(1:m) = WK(1:m)
  od_layer%wbroad = WBROAD
  od_layer%dv = DV
  od_layer%v1 = V1
  od_layer%v2 = V2
  od_layer%nmol = int(NMOL, c_int32_t)
  od_layer%jrad = int(JRAD, c_int32_t)
  rc = lblx_od_size(r, od_layer, n)
  if (rc /= LBLX_OK) call lblx_stop('OPDPTH', rc)
  if (allocated(od)) deallocate (od)
  allocate (od(n))
  rc = lblx_od_layer(ctx, r, od_layer, od, n, v1p0, c_null_ptr)
  if (rc /= LBLX_OK) call lblx_stop('OPDPTH', rc)
  JRAD = od_layer%jrad                     ! FSCDID(9), as OPDPTH leaves it
  od_run = r
  od_valid = .true.
  ! HIRAC1 with IOD = 0: the file header, then PANEL's records
  rc = lblx_od_panels(r, od_layer, 0_c_int, q1, q1, qn, np)    ! the panel count
  allocate (v1p(np), v2p(np), nlim(np))
  rc = lblx_od_panels(r, od_layer, np, v1p, v2p, nlim, np)
  if (rc /= LBLX_OK) call lblx_stop('OPDPTH', rc)
  call lblx_bufout(KFILE, FILHDR, NFHDRF)
  DVP = DV
  o = 0
  do k = 1, np
     V1PK = v1p(k)
     V2PK = v2p(k)
     NLIMK = nlim(k)
     write (KFILE) V1PK, V2PK, DVP, NLIMK
     write (KFILE) (real(od(o + i)), i = 1, NLIMK)
     o = o + NLIMK
  end do
  deallocate (v1p, v2p, nlim)
END SUBROUTINE OPDPTH

SUBROUTINE XMERGE (NPTS,LFILE,MFILE,JPATHL)
  use lblx_shim_state
  IMPLICIT REAL*8 (V)
  COMMON /IFIL/ IRD,IPR,IPU,NOPR,NFHDRF,NPHDRF,NFHDRL,NPHDRL,          &
       &        NLNGTH,KFILE,KPANEL,LINFIL,NFILE,IAFIL,IEXFIL,           &
       &        NLTEFL,LNFIL4,LNGTH4,IBRD
  COMMON /MANE/ P0,TEMP0,NLAYER,DVXM,H2OSLF,WTOT,ALBAR,ADBAR,AVBAR,    &
       &        AVFIX,LAYER,SECNT0,SAMPLE,DVSET,ALFAL0,AVMASS,           &
       &        DPTMIN,DPTFAC,ALTAV,AVTRAT,TDIFF1,TDIFF2,ALTD1,          &
       &        ALTD2,ANGLE,IANT,LTGNT,LH1,LH2,IPFLAG,PLAY,TLAY,         &
       &        EXTID(10)
  CHARACTER*8 EXTID
  character*8 XID, HMOLID, YID
  real*8 SECANT, XALTZ
  COMMON /FILHDR/ XID(10),SECANT,PAVE,TAVE,HMOLID(60),XALTZ(4),        &
       &          WK(60),PZL,PZU,TZL,TZU,WBROAD,DV ,V1 ,V2 ,TBOUND,      &
       &          EMISIV,FSCDID(17),NMOL,LAYHDR,YI1,YID(10),LSTWDF
  COMMON /BNDPRP/ TMPBND,BNDEMI(3),BNDRFL(3),IBPROP,surf_refl,         &
       &          pad_3,angle_path, secant_diffuse, secant_path, diffuse_fac
  character*1 surf_refl
  character*3 pad_3
  DIMENSION FILHDR(2)
  EQUIVALENCE (FSCDID(1),IHIRAC) , (FSCDID(2),ILBLF4),                 &
       &      (FSCDID(3),IXSCNT) , (FSCDID(4),IAERSL),                   &
       &      (FSCDID(5),IEMIT) , (FSCDID(8),IPATHL),                    &
       &      (FSCDID(9),JRAD) , (FSCDID(11),IMRG),                      &
       &      (FSCDID(16),LAYR1) , (XID(1),FILHDR(1))
  type(lblx_layer), target :: lay
  integer(c_int) :: rc, np
  integer(c_int64_t) :: n, nb
  type(c_ptr) :: d_od, d_rnew, d_tnew, pold
  real(c_double), allocatable, target :: rad(:), tra(:)
  real(c_double), allocatable :: v1p(:), v2p(:)
  integer(c_int32_t), allocatable :: nlim(:)
  real(c_double) :: tbnd, q1(1)
  integer(c_int32_t) :: qn(1)
  integer :: k, i, o
  REAL*8 V1PK, V2PK
  REAL DVP
  INTEGER NLIMK

  IF (JPATHL.GE.1) IPATHL = JPATHL
  if (IEMIT /= 1 .or. (IPATHL /= 1 .and. IPATHL /= 3) .or. IMRG == 36 .or. IMRG == 46 .or. IMRG == 10 .or. &
       (IAERSL >= 1 .and. IAERSL /= 5) .or. IBPROP == 1 .or. BNDEMI(2) /= 0. .or. BNDEMI(3) /= 0.) &
       stop 'XMERGE: option outside the liblblx path (IEMIT, IPATHL, IMRG, aerosols, surface)'
  if (.not. od_valid .or. od_layer%dv /= DV .or. od_layer%v1 /= V1 .or. od_layer%v2 /= V2 .or. &
       od_layer%pave /= PAVE .or. od_layer%tave /= TAVE) &
       stop 'XMERGE: the layer optical depth of this layer is not the last OPDPTH result'
  lay = od_layer
  tbnd = 0.
  pold = c_null_ptr
  if (LAYER == LH1) then                     ! EMINIT: the layer at the surface
     if (IPATHL == 1) tbnd = TMPBND          ! (observer's layer for IPATHL = 3)
     old_valid = .false.
  else                                       ! RADMRG onto the merged layers so far
     if (.not. old_valid) stop 'XMERGE: no merged result of the layers below'
     if (IPATHL == 3 .and. LAYER == LH2) tbnd = TMPBND   ! EMBND (src/xmerge.f90:140)
     pold = c_loc(old_layer)
  end if
  rc = lblx_merge_size(od_run, pold, lay, n)
  if (rc /= LBLX_OK) call lblx_stop('XMERGE', rc)
  nb = 8_c_int64_t * size(od, kind=c_int64_t)
  d_od = lblx_dev_alloc(ctx, nb)
  d_rnew = lblx_dev_alloc(ctx, 8_c_int64_t * n)
  d_tnew = lblx_dev_alloc(ctx, 8_c_int64_t * n)
  if (.not. (c_associated(d_od) .and. c_associated(d_rnew) .and. c_associated(d_tnew))) &
       stop 'XMERGE: device memory'
  rc = lblx_copy(ctx, d_od, c_loc(od), nb, 0_c_int)
  if (rc /= LBLX_OK) call lblx_stop('XMERGE', rc)
  if (IPATHL == 3) then                      ! EMIN at the observer's side, TZL
     rc = lblx_xmerge_path_device(ctx, od_run, pold, lay, 3_c_int, TZL, tbnd, BNDEMI(1), d_od, d_rold, &
          d_told, d_rnew, d_tnew)
  else
     rc = lblx_xmerge_device(ctx, od_run, pold, lay, TZU, tbnd, BNDEMI(1), d_od, d_rold, d_told, d_rnew, d_tnew)
  end if
  if (rc /= LBLX_OK) call lblx_stop('XMERGE', rc)
  allocate (rad(n), tra(n))
  rc = lblx_copy(ctx, c_loc(rad), d_rnew, 8_c_int64_t * n, 1_c_int)
  if (rc == LBLX_OK) rc = lblx_copy(ctx, c_loc(tra), d_tnew, 8_c_int64_t * n, 1_c_int)
  if (rc /= LBLX_OK) call lblx_stop('XMERGE', rc)
  rc = lblx_dev_free(ctx, d_od)
  if (c_associated(d_rold)) rc = lblx_dev_free(ctx, d_rold)
  if (c_associated(d_told)) rc = lblx_dev_free(ctx, d_told)
  d_rold = d_rnew
  d_told = d_tnew
  old_layer = lay
  old_valid = .true.
  ! EMINIT/RADMRG: the file header with IEMIT = 1 and LAYR1 = LAYER, then
  ! EMOUT's records for each panel of the layer (src/xmerge.f90:2394-2401,
  ! 2862, 3737-3790)
  IEMIT = 1
  LAYR1 = LAYER
  call lblx_bufout(MFILE, FILHDR, NFHDRF)
  rc = lblx_od_panels(od_run, lay, 0_c_int, q1, q1, qn, np)      ! the panel count
  allocate (v1p(np), v2p(np), nlim(np))
  rc = lblx_od_panels(od_run, lay, np, v1p, v2p, nlim, np)
  if (rc /= LBLX_OK) call lblx_stop('XMERGE', rc)
  DVP = DV
  o = 0
  do k = 1, np
     V1PK = v1p(k)
     V2PK = v2p(k)
     NLIMK = nlim(k)
     write (MFILE) V1PK, V2PK, DVP, NLIMK
     write (MFILE) (real(rad(o + i)), i = 1, NLIMK)
     write (MFILE) (real(tra(o + i)), i = 1, NLIMK)
     o = o + NLIMK
  end do
  deallocate (v1p, v2p, nlim, rad, tra)
END SUBROUTINE XMERGE
