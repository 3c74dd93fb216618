! lblx_c -- Fortran (ISO_C_BINDING) interface to the C ABI of liblblx
! (include/lblx.h): the bind(c) mirrors of its structures and interfaces to
! its entry points, for the reference's Fortran driver.  Every structure
! component has an explicit C kind, so the layout is the same whatever the
! default INTEGER/REAL kinds of the build (-fdefault-integer-8
! -fdefault-real-8 for the double-precision build).
module lblx_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: LBLX_OK = 0, LBLX_E_ARG = -1, LBLX_E_NODEV = -2, LBLX_E_HIP = -3, &
       LBLX_E_NREC = -4, LBLX_E_UNSUPPORTED = -5, LBLX_E_SIZE = -6, LBLX_E_NOLINES = -7, &
       LBLX_E_DVRATIO = -8, LBLX_E_COARSER = -9, LBLX_E_SCRATCH = -10
  integer(c_int32_t), parameter :: LBLX_FP64 = 0, LBLX_FP32_SPLIT = 1
  integer, parameter :: LBLX_MXMOL = 64

  type, bind(c) :: lblx_run
     real(c_double) :: v1, v2, sample, alfal0, avmass, dptmin, dptfac
     integer(c_int32_t) :: ilblf4, icntnm, irad, precision
     real(c_double) :: cntscl(7)
     integer(c_int32_t) :: ifrgnx, iaersl, voigt, ibrd
  end type lblx_run

  type, bind(c) :: lblx_layer
     real(c_double) :: pave, tave, wk(LBLX_MXMOL), wbroad, dv, v1, v2
     integer(c_int32_t) :: nmol, jrad
  end type lblx_layer

  type, bind(c) :: lblx_stats
     integer(c_int64_t) :: nlin, lincnt, nlins4, ilin4, npanel, npts, k1_updates, k3_updates, nwin4, &
          nfix, fixseq
  end type lblx_stats

  type, bind(c) :: lblx_scan
     real(c_double) :: hwhm, v1, v2, dvout, param
     integer(c_int32_t) :: jemit, jfnin, mratin, ivx, nofix, pad
  end type lblx_scan

  type, bind(c) :: lblx_scanfn
     real(c_double) :: hwhm, v1, v2, sampl
     integer(c_int32_t) :: jemit, jfn, jvar, npanel
  end type lblx_scanfn

  interface
     integer(c_int) function lblx_create(ctx, device) bind(c, name="lblx_create")
       import :: c_int, c_ptr
       type(c_ptr) :: ctx
       integer(c_int), value :: device
     end function lblx_create
     subroutine lblx_destroy(ctx) bind(c, name="lblx_destroy")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine lblx_destroy
     type(c_ptr) function lblx_strerror(code) bind(c, name="lblx_strerror")
       import :: c_int, c_ptr
       integer(c_int), value :: code
     end function lblx_strerror
     integer(c_int) function lblx_device_count() bind(c, name="lblx_device_count")
       import :: c_int
     end function lblx_device_count
     integer(c_int) function lblx_load_tape3(ctx, path) bind(c, name="lblx_load_tape3")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char) :: path(*)
     end function lblx_load_tape3
     integer(c_int) function lblx_plan_dv(run, nlay, nmol, pave, tave, wk, wbroad, iemit, dv, itype) &
          bind(c, name="lblx_plan_dv")
       import :: c_int, c_double, c_int32_t, lblx_run
       type(lblx_run), intent(in) :: run
       integer(c_int), value :: nlay, nmol, iemit
       real(c_double), intent(in) :: pave(*), tave(*), wk(*), wbroad(*)
       real(c_double) :: dv(*)
       integer(c_int32_t) :: itype(*)
     end function lblx_plan_dv
     integer(c_int) function lblx_od_size(run, layer, npts) bind(c, name="lblx_od_size")
       import :: c_int, c_int64_t, lblx_run, lblx_layer
       type(lblx_run), intent(in) :: run
       type(lblx_layer), intent(in) :: layer
       integer(c_int64_t) :: npts
     end function lblx_od_size
     integer(c_int) function lblx_od_panels(run, layer, cap, v1p, v2p, nlim, npanel) &
          bind(c, name="lblx_od_panels")
       import :: c_int, c_double, c_int32_t, lblx_run, lblx_layer
       type(lblx_run), intent(in) :: run
       type(lblx_layer), intent(in) :: layer
       integer(c_int), value :: cap
       real(c_double) :: v1p(*), v2p(*)
       integer(c_int32_t) :: nlim(*)
       integer(c_int) :: npanel
     end function lblx_od_panels
     integer(c_int) function lblx_od_layer(ctx, run, layer, od, n_out, v1p, st) bind(c, name="lblx_od_layer")
       import :: c_int, c_ptr, c_double, c_int64_t, lblx_run, lblx_layer
       type(c_ptr), value :: ctx
       type(lblx_run), intent(in) :: run
       type(lblx_layer) :: layer
       real(c_double) :: od(*)
       integer(c_int64_t), value :: n_out
       real(c_double) :: v1p
       type(c_ptr), value :: st
     end function lblx_od_layer
     integer(c_int) function lblx_merge_size(run, layer_old, layer, n) bind(c, name="lblx_merge_size")
       import :: c_int, c_ptr, c_int64_t, lblx_run, lblx_layer
       type(lblx_run), intent(in) :: run
       type(c_ptr), value :: layer_old
       type(lblx_layer), intent(in) :: layer
       integer(c_int64_t) :: n
     end function lblx_merge_size
     integer(c_int) function lblx_xmerge_device(ctx, run, layer_old, layer, tzu, tbnd, emis, d_od, d_rold, &
          d_told, d_rnew, d_tnew) bind(c, name="lblx_xmerge_device")
       import :: c_int, c_ptr, c_double, lblx_run, lblx_layer
       type(c_ptr), value :: ctx
       type(lblx_run), intent(in) :: run
       type(c_ptr), value :: layer_old
       type(lblx_layer), intent(in) :: layer
       real(c_double), value :: tzu, tbnd, emis
       type(c_ptr), value :: d_od, d_rold, d_told, d_rnew, d_tnew
     end function lblx_xmerge_device
     integer(c_int) function lblx_xmerge_path_device(ctx, run, layer_old, layer, ipathl, tz, tbnd, emis, &
          d_od, d_rold, d_told, d_rnew, d_tnew) bind(c, name="lblx_xmerge_path_device")
       import :: c_int, c_ptr, c_double, lblx_run, lblx_layer
       type(c_ptr), value :: ctx
       type(lblx_run), intent(in) :: run
       type(c_ptr), value :: layer_old
       type(lblx_layer), intent(in) :: layer
       integer(c_int), value :: ipathl
       real(c_double), value :: tz, tbnd, emis
       type(c_ptr), value :: d_od, d_rold, d_told, d_rnew, d_tnew
     end function lblx_xmerge_path_device
     integer(c_int) function lblx_scan_size(scan, v1_in, dv_in, n_in, n_out, v1_out, dv_out) &
          bind(c, name="lblx_scan_size")
       import :: c_int, c_double, c_int64_t, lblx_scan
       type(lblx_scan), intent(in) :: scan
       real(c_double), value :: v1_in, dv_in
       integer(c_int64_t), value :: n_in
       integer(c_int64_t) :: n_out
       real(c_double) :: v1_out, dv_out
     end function lblx_scan_size
     integer(c_int) function lblx_scanfn_size(scan, iemit, v1_in, dv_in, n_in, n_out, v1_out, dv_out) &
          bind(c, name="lblx_scanfn_size")
       import :: c_int, c_double, c_int64_t, lblx_scanfn
       type(lblx_scanfn), intent(in) :: scan
       integer(c_int), value :: iemit
       real(c_double), value :: v1_in, dv_in
       integer(c_int64_t), value :: n_in
       integer(c_int64_t) :: n_out
       real(c_double) :: v1_out, dv_out
     end function lblx_scanfn_size
     integer(c_int) function lblx_scanfn_device(ctx, scan, iemit, d_in, v1_in, dv_in, n_in, d_out, n_out) &
          bind(c, name="lblx_scanfn_device")
       import :: c_int, c_ptr, c_double, c_int64_t, lblx_scanfn
       type(c_ptr), value :: ctx
       type(lblx_scanfn), intent(in) :: scan
       integer(c_int), value :: iemit
       type(c_ptr), value :: d_in, d_out
       real(c_double), value :: v1_in, dv_in
       integer(c_int64_t), value :: n_in, n_out
     end function lblx_scanfn_device
     integer(c_int) function lblx_fftscn_device(ctx, scan, d_in, v1_in, dv_in, n_in, d_out, cap, n_out, &
          v1_out, dv_out) bind(c, name="lblx_fftscn_device")
       import :: c_int, c_ptr, c_double, c_int64_t, lblx_scan
       type(c_ptr), value :: ctx
       type(lblx_scan), intent(in) :: scan
       type(c_ptr), value :: d_in, d_out
       real(c_double), value :: v1_in, dv_in
       integer(c_int64_t), value :: n_in, cap
       integer(c_int64_t) :: n_out
       real(c_double) :: v1_out, dv_out
     end function lblx_fftscn_device
     integer(c_int) function lblx_sync(ctx) bind(c, name="lblx_sync")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function lblx_sync
     type(c_ptr) function lblx_dev_alloc(ctx, nbytes) bind(c, name="lblx_dev_alloc")
       import :: c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: nbytes
     end function lblx_dev_alloc
     integer(c_int) function lblx_dev_free(ctx, p) bind(c, name="lblx_dev_free")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, p
     end function lblx_dev_free
     integer(c_int) function lblx_copy(ctx, dst, src, nbytes, kind) bind(c, name="lblx_copy")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx, dst, src
       integer(c_int64_t), value :: nbytes
       integer(c_int), value :: kind
     end function lblx_copy
  end interface

contains

  ! The message of a return code as a Fortran string.
  function lblx_message(code) result(msg)
    integer(c_int), intent(in) :: code
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    p = lblx_strerror(code)
    msg = ''
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
       if (s(n + 1) == c_null_char) exit
       n = n + 1
    end do
    allocate (character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function lblx_message

end module lblx_c
