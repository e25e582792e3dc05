! Held-Suarez driver with the Fortran host in charge: sets up the T42L20
! (or T254L64) configuration, runs run.f90's time loop on the GPU through
! libgfsdc and prints run.f90's per-step diagnostic line (run.f90:159-163).
! Also exercises module shtns (the transform wrapper) with a round trip and a
! solid-body rotation through getuv.
! With a fourth argument ntdfi > 0 the run starts with run.f90's digital
! filter window of 2 ntdfi steps (run.f90:58-145) and continues from the
! window's middle.
! Usage: gfsdc_hs_driver [nsteps] [T42L20|T254L64] [fp64|fp32] [ntdfi]
!        gfsdc_hs_driver abi   (layout check of gfsdc_cfg only; no GPU call)
program hs_driver
  use iso_c_binding
  use kinds, only: r_kind
  use shtns, only: shtns_init, spectogrd, grdtospec, getuv, nlm, lats
  use gfsdc_dycore
  implicit none
  type(gfsdc_cfg) :: cfg
  type(c_ptr) :: h
  integer :: nsteps, nt, k, nlev, ntdfi
  character(len=16) :: arg, cname, prec
  real(c_double), allocatable, target :: ak(:), bk(:), diag(:)
  complex(r_kind), allocatable :: spec(:), spec2(:), zero(:)
  real(r_kind), allocatable :: grid(:,:), ug(:,:), vg(:,:)
  real(c_double) :: t, eta, b, err, u0, pi
  real :: tc0, tc1
  nsteps = 72
  cname = 'T42L20'
  prec = 'fp64'
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg)
    if (arg == 'abi') then
      call dyc_check_abi()
      print '(a,i0,a,i0)', 'gfsdc_cfg bytes: fortran ', c_sizeof(cfg), ' library ', dyc_cfg_size()
      stop
    end if
    read (arg, *) nsteps
  end if
  if (command_argument_count() >= 2) call get_command_argument(2, cname)
  if (command_argument_count() >= 3) call get_command_argument(3, prec)
  ntdfi = 0
  if (command_argument_count() >= 4) then
    call get_command_argument(4, arg)
    read (arg, *) ntdfi
  end if
  call dyc_defaults(cfg)
  if (cname == 'T254L64') then
    cfg%nlon = 768; cfg%nlat = 384; cfg%ntrunc = 254; cfg%nlev = 64; cfg%dt = 600.d0
  else
    cfg%nlon = 128; cfg%nlat = 64; cfg%ntrunc = 42; cfg%nlev = 20; cfg%dt = 1200.d0
  end if
  nlev = cfg%nlev
  allocate(ak(nlev + 1), bk(nlev + 1), diag(4))
  do k = 0, nlev
    eta = real(k, c_double) / nlev
    if (nlev == 20) then          ! equally spaced sigma
      ak(k + 1) = 0.d0; bk(k + 1) = eta
    else                          ! hybrid: B = ((eta-0.2)/0.8)**1.5, A = p0 (eta - B)
      b = max(0.d0, (eta - 0.2d0) / 0.8d0) ** 1.5d0
      if (k == nlev) b = 1.d0
      bk(k + 1) = b; ak(k + 1) = 1.d5 * (eta - b)
      if (k == nlev) ak(k + 1) = 0.d0
    end if
  end do
  cfg%ak = c_loc(ak)
  cfg%bk = c_loc(bk)
  cfg%physics = PHYS_HS
  cfg%precision = GFSDC_FP64
  if (prec == 'fp32') cfg%precision = GFSDC_FP32
  if (dyc_create(cfg, h) /= 0) stop 'gfsdc_create failed'
  if (dyc_init_rest_state(h, 300.d0, 1.d5, 1.d-6, 1, 21, 0_c_long_long) /= 0) stop 'init failed'
  ! transform wrapper round trip on the same grid (module shtns)
  call shtns_init(cfg%nlon, cfg%nlat, cfg%ntrunc)
  allocate(spec(nlm), spec2(nlm), grid(cfg%nlon, cfg%nlat))
  spec = 0
  spec(2) = (1.0, 0.0)
  spec(cfg%ntrunc + 3) = (0.5, -0.25)
  call spectogrd(spec, grid)
  call grdtospec(grid, spec2)
  err = maxval(abs(spec2 - spec))
  print '(a,es10.3)', 'shtns round trip max error = ', err
  ! solid-body rotation: vorticity 2 u0 sin(lat) / a, i.e. q_10 = 2 u0 / (a sqrt(3/4pi)),
  ! gives u = u0 cos(lat), v = 0
  pi = 4.d0 * atan(1.d0)
  u0 = 20.d0
  allocate(zero(nlm), ug(cfg%nlon, cfg%nlat), vg(cfg%nlon, cfg%nlat))
  zero = 0
  spec = 0
  spec(2) = cmplx(2.d0 * u0 / (cfg%rerth * sqrt(3.d0 / (4.d0 * pi))), 0.d0, r_kind)
  call getuv(spec, zero, ug, vg, real(cfg%rerth, r_kind))
  err = max(maxval(abs(ug - u0 * cos(lats))), maxval(abs(vg))) / u0
  print '(a,es10.3)', 'shtns getuv solid-body max error = ', err
  if (ntdfi > 0) then
    ! run.f90:58-145: the filter window on the device, then the state reset to its
    ! weighted average and the clock to the window's middle
    if (dyc_dfi(h, ntdfi, 1) /= 0) stop 'dfi failed'
    if (dyc_get_time(h, t) /= 0) stop 'clock failed'
    print '(a,f0.3,a)', 'done with dfi loop, restarting at t = ', t / 3600.d0, ' hrs'
  end if
  do nt = 1, nsteps
    call cpu_time(tc0)
    if (dyc_advance(h, 1, 1) /= 0) stop 'advance failed'
    if (dyc_step_diag(h, c_loc(diag)) /= 0) stop 'diag failed'
    call cpu_time(tc1)
    if (dyc_get_time(h, t) /= 0) stop 'clock failed'
    write (6, 8998) t / 3600.d0, diag(1), diag(2) / 100., diag(3) / 100., diag(4), tc1 - tc0
  end do
8998 format('t = ',f0.3,' hrs, spdmax = ',f7.3,', min/max ps = ',f7.2,'/',f7.2,', pstend = ',f0.3,', cpu time = ',f0.3)
  if (dyc_destroy(h) /= 0) stop 'destroy failed'
end program hs_driver
