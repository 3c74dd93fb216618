! Test driver: the layout of the bind(c) types of lblx_c as flang lays
! them out (offset and size of every component, size of every type), and
! the host-only entry points called through the Fortran interfaces.  The
! output is compared with the C compiler's view of include/lblx.h.
program abi_layout
  use iso_c_binding
  use lblx_c
  implicit none
  type(lblx_run), target :: r
  type(lblx_layer), target :: l
  type(lblx_stats), target :: s
  type(lblx_scan), target :: c
  type(lblx_scanfn), target :: sf
  real(c_double) :: pave(1), tave(1), wk(1), wb(1), dv(1), q1(1)
  integer(c_int32_t) :: it(1), qn(1)
  integer(c_int64_t) :: n
  integer(c_int) :: rc, np

  call f('lblx_run', 'v1', c_loc(r), c_loc(r%v1), c_sizeof(r%v1))
  call f('lblx_run', 'v2', c_loc(r), c_loc(r%v2), c_sizeof(r%v2))
  call f('lblx_run', 'sample', c_loc(r), c_loc(r%sample), c_sizeof(r%sample))
  call f('lblx_run', 'alfal0', c_loc(r), c_loc(r%alfal0), c_sizeof(r%alfal0))
  call f('lblx_run', 'avmass', c_loc(r), c_loc(r%avmass), c_sizeof(r%avmass))
  call f('lblx_run', 'dptmin', c_loc(r), c_loc(r%dptmin), c_sizeof(r%dptmin))
  call f('lblx_run', 'dptfac', c_loc(r), c_loc(r%dptfac), c_sizeof(r%dptfac))
  call f('lblx_run', 'ilblf4', c_loc(r), c_loc(r%ilblf4), c_sizeof(r%ilblf4))
  call f('lblx_run', 'icntnm', c_loc(r), c_loc(r%icntnm), c_sizeof(r%icntnm))
  call f('lblx_run', 'irad', c_loc(r), c_loc(r%irad), c_sizeof(r%irad))
  call f('lblx_run', 'precision', c_loc(r), c_loc(r%precision), c_sizeof(r%precision))
  call f('lblx_run', 'cntscl', c_loc(r), c_loc(r%cntscl), c_sizeof(r%cntscl))
  call f('lblx_run', 'ifrgnx', c_loc(r), c_loc(r%ifrgnx), c_sizeof(r%ifrgnx))
  call f('lblx_run', 'iaersl', c_loc(r), c_loc(r%iaersl), c_sizeof(r%iaersl))
  call f('lblx_run', 'voigt', c_loc(r), c_loc(r%voigt), c_sizeof(r%voigt))
  call f('lblx_run', 'ibrd', c_loc(r), c_loc(r%ibrd), c_sizeof(r%ibrd))
  write (*, '(a,1x,a,1x,i0)') 'size', 'lblx_run', c_sizeof(r)
  call f('lblx_layer', 'pave', c_loc(l), c_loc(l%pave), c_sizeof(l%pave))
  call f('lblx_layer', 'tave', c_loc(l), c_loc(l%tave), c_sizeof(l%tave))
  call f('lblx_layer', 'wk', c_loc(l), c_loc(l%wk), c_sizeof(l%wk))
  call f('lblx_layer', 'wbroad', c_loc(l), c_loc(l%wbroad), c_sizeof(l%wbroad))
  call f('lblx_layer', 'dv', c_loc(l), c_loc(l%dv), c_sizeof(l%dv))
  call f('lblx_layer', 'v1', c_loc(l), c_loc(l%v1), c_sizeof(l%v1))
  call f('lblx_layer', 'v2', c_loc(l), c_loc(l%v2), c_sizeof(l%v2))
  call f('lblx_layer', 'nmol', c_loc(l), c_loc(l%nmol), c_sizeof(l%nmol))
  call f('lblx_layer', 'jrad', c_loc(l), c_loc(l%jrad), c_sizeof(l%jrad))
  write (*, '(a,1x,a,1x,i0)') 'size', 'lblx_layer', c_sizeof(l)
  call f('lblx_stats', 'nlin', c_loc(s), c_loc(s%nlin), c_sizeof(s%nlin))
  call f('lblx_stats', 'npts', c_loc(s), c_loc(s%npts), c_sizeof(s%npts))
  call f('lblx_stats', 'nwin4', c_loc(s), c_loc(s%nwin4), c_sizeof(s%nwin4))
  call f('lblx_stats', 'nfix', c_loc(s), c_loc(s%nfix), c_sizeof(s%nfix))
  call f('lblx_stats', 'fixseq', c_loc(s), c_loc(s%fixseq), c_sizeof(s%fixseq))
  write (*, '(a,1x,a,1x,i0)') 'size', 'lblx_stats', c_sizeof(s)
  call f('lblx_scan', 'hwhm', c_loc(c), c_loc(c%hwhm), c_sizeof(c%hwhm))
  call f('lblx_scan', 'param', c_loc(c), c_loc(c%param), c_sizeof(c%param))
  call f('lblx_scan', 'jemit', c_loc(c), c_loc(c%jemit), c_sizeof(c%jemit))
  call f('lblx_scan', 'nofix', c_loc(c), c_loc(c%nofix), c_sizeof(c%nofix))
  write (*, '(a,1x,a,1x,i0)') 'size', 'lblx_scan', c_sizeof(c)
  call f('lblx_scanfn', 'hwhm', c_loc(sf), c_loc(sf%hwhm), c_sizeof(sf%hwhm))
  call f('lblx_scanfn', 'sampl', c_loc(sf), c_loc(sf%sampl), c_sizeof(sf%sampl))
  call f('lblx_scanfn', 'jemit', c_loc(sf), c_loc(sf%jemit), c_sizeof(sf%jemit))
  call f('lblx_scanfn', 'npanel', c_loc(sf), c_loc(sf%npanel), c_sizeof(sf%npanel))
  write (*, '(a,1x,a,1x,i0)') 'size', 'lblx_scanfn', c_sizeof(sf)

  ! the single-layer workload's run and layer through the host-only calls
  r = lblx_run(600.d0, 1100.d0, 4.d0, 0.04d0, 36.d0, 2.d-4, 1.d-3, 1, 0, 0, LBLX_FP64, &
       [0.d0, 0.d0, 0.d0, 0.d0, 0.d0, 0.d0, 0.d0], 0, 0, 0, 0)
  pave(1) = 500.d0
  tave(1) = 250.d0
  wk(1) = 1.d22
  wb(1) = 1.448766d24
  rc = lblx_plan_dv(r, 1_c_int, 1_c_int, pave, tave, wk, wb, 0_c_int, dv, it)
  write (*, '(a,1x,i0,1x,es24.16e3)') 'plan_dv', rc, dv(1)
  l%pave = pave(1)
  l%tave = tave(1)
  l%wk = 0.d0
  l%wk(1) = wk(1)
  l%wbroad = wb(1)
  l%dv = dv(1)
  l%v1 = 600.d0
  l%v2 = 1100.d0
  l%nmol = 1
  l%jrad = 1
  rc = lblx_od_size(r, l, n)
  write (*, '(a,1x,i0,1x,i0)') 'od_size', rc, n
  rc = lblx_od_panels(r, l, 0_c_int, q1, q1, qn, np)
  write (*, '(a,1x,i0,1x,i0)') 'od_panels', rc, np
  write (*, '(a,1x,a)') 'strerror', lblx_message(LBLX_E_UNSUPPORTED)

contains
  subroutine f(ty, name, base, field, sz)
    character(len=*), intent(in) :: ty, name
    type(c_ptr), intent(in) :: base, field
    integer(c_size_t), intent(in) :: sz
    integer(c_intptr_t) :: a, b
    a = transfer(base, a)
    b = transfer(field, b)
    write (*, '(a,1x,a,1x,a,1x,i0,1x,i0)') 'off', ty, name, b - a, sz
  end subroutine f
end program abi_layout
