! module gfsdc_dycore: iso_c_binding interface of the device-resident
! dycore tier of libgfsdc (include/gfsdc.h). A Fortran host keeps the
! reference's call sequence (run.f90 advance -> dyn_run.f90 getdyntend ->
! semi-implicit solve -> damping -> physics) and calls these entry points
! instead of looping over levels on the host.
module gfsdc_dycore
  use iso_c_binding
  implicit none
  private
  public :: gfsdc_cfg, dyc_defaults, dyc_cfg_size, dyc_check_abi, dyc_create, dyc_destroy, dyc_set_state, dyc_get_state, &
            dyc_getdyntend, dyc_advance, dyc_step_diag, dyc_init_rest_state, dyc_sync, &
            dyc_set_time, dyc_get_time, dyc_dfi, dyc_iau_set, dyc_checkpoint_size, dyc_checkpoint_save, &
            dyc_checkpoint_load, gfsdc_stoch_cfg, dyc_stoch_defaults, dyc_stoch_init, dyc_get_stoch, &
            dyc_mlp_init, dyc_mlp_collect, dyc_mlp_normalize, dyc_mlp_train, dyc_mlp_enable, &
            dyc_mlp_judge_stats, dyc_mlp_params, dyc_mlp_norm, &
            GFSDC_FP64, GFSDC_FP32, GFSDC_HOST, DT_F64, DT_F32, PHYS_NONE, PHYS_HS, PHYS_SIMPLE

  integer(c_int), parameter :: GFSDC_FP64 = 0, GFSDC_FP32 = 1, GFSDC_HOST = 0, DT_F64 = 0, DT_F32 = 1
  integer(c_int), parameter :: PHYS_NONE = 0, PHYS_HS = 1, PHYS_SIMPLE = 2

  type, bind(c) :: gfsdc_cfg
    integer(c_int) :: nlon, nlat, ntrunc, nlev, ntrac, nmem, precision
    real(c_double) :: dt
    type(c_ptr) :: ak, bk
    integer(c_int) :: explicit_, physics, massfix, ndiss
    real(c_double) :: efold, hdif_fac, hdif_fac2, fshk, taustratdamp
    real(c_double) :: rerth, grav, omega, rd, rv, cp, cvap
    integer(c_int) :: device
    integer(c_int) :: simple_test, turbflux
    real(c_double) :: tpat_amp, tpat_tau, tpat_lscale
    integer(c_int) :: tpat_seed
  end type

  ! stochastic physics settings (params.f90:140-160 names, include/gfsdc.h)
  type, bind(c) :: gfsdc_stoch_cfg
    real(c_double) :: vcamp
    real(c_double) :: svc, svc_tau, svc_lscale
    integer(c_int) :: iseed_svc
    real(c_double) :: sppt, sppt_tau, sppt_lscale
    integer(c_int) :: iseed_sppt
    real(c_double) :: shum, shum_tau, shum_lscale
    integer(c_int) :: iseed_shum
    real(c_double) :: addnoise, addnoise_tau, addnoise_lscale
    integer(c_int) :: iseed_addnoise
    integer(c_int) :: addnoise_vfilt, addnoise_kenorm, addnoise_dissfact, addnoise_vrtonly
  end type

  interface
    ! sizeof(gfsdc_cfg) in the library: the layout check of dyc_check_abi
    integer(c_size_t) function dyc_cfg_size() bind(c, name="gfsdc_cfg_size")
      import :: c_size_t
    end function
    subroutine dyc_defaults(cfg) bind(c, name="gfsdc_cfg_defaults")
      import :: gfsdc_cfg
      type(gfsdc_cfg), intent(inout) :: cfg
    end subroutine
    integer(c_int) function dyc_create(cfg, out) bind(c, name="gfsdc_create")
      import :: gfsdc_cfg, c_ptr, c_int
      type(gfsdc_cfg), intent(in) :: cfg
      type(c_ptr), intent(out) :: out
    end function
    integer(c_int) function dyc_destroy(h) bind(c, name="gfsdc_destroy")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
    end function
    integer(c_int) function dyc_set_state(h, vrt, div, tv, trac, lnps, dtype, loc) bind(c, name="gfsdc_set_state")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, vrt, div, tv, trac, lnps
      integer(c_int), value :: dtype, loc
    end function
    integer(c_int) function dyc_get_state(h, vrt, div, tv, trac, lnps, dtype, loc) bind(c, name="gfsdc_get_state")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, vrt, div, tv, trac, lnps
      integer(c_int), value :: dtype, loc
    end function
    integer(c_int) function dyc_getdyntend(h, stage, inv, dvrt, ddiv, dtv, dtrac, dlnps, dtype, loc) &
        bind(c, name="gfsdc_getdyntend")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, dvrt, ddiv, dtv, dtrac, dlnps
      integer(c_int), value :: stage, inv, dtype, loc
    end function
    integer(c_int) function dyc_advance(h, nsteps, use_graph) bind(c, name="gfsdc_advance")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps, use_graph
    end function
    integer(c_int) function dyc_step_diag(h, out) bind(c, name="gfsdc_step_diag")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, out
    end function
    integer(c_int) function dyc_init_rest_state(h, t0, ps0, vrms, nmin, nmax, seed) &
        bind(c, name="gfsdc_init_rest_state")
      import :: c_ptr, c_int, c_double, c_long_long
      type(c_ptr), value :: h
      real(c_double), value :: t0, ps0, vrms
      integer(c_int), value :: nmin, nmax
      integer(c_long_long), value :: seed
    end function
    integer(c_int) function dyc_sync(h) bind(c, name="gfsdc_sync")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
    end function
    ! model clock (s): run.f90:53 t = tstart, t = t + dt per step
    integer(c_int) function dyc_set_time(h, t) bind(c, name="gfsdc_set_time")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), value :: t
    end function
    integer(c_int) function dyc_get_time(h, t) bind(c, name="gfsdc_get_time")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: t
    end function
    ! run.f90:58-145 digital filter initialisation (2 ntdfi steps)
    integer(c_int) function dyc_dfi(h, ntdfi, use_graph) bind(c, name="gfsdc_dfi")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: ntdfi, use_graph
    end function
    ! iau_module.f90 init_iau with the increments formed by the caller
    integer(c_int) function dyc_iau_set(h, nfiles, fhrs, delthrs, vrt, div, tv, trac, lnps, dtype, loc) &
        bind(c, name="gfsdc_iau_set")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h, fhrs, vrt, div, tv, trac, lnps
      integer(c_int), value :: nfiles, dtype, loc
      real(c_double), value :: delthrs
    end function
    ! device checkpoint / resume
    integer(c_int) function dyc_checkpoint_size(h, nbytes) bind(c, name="gfsdc_checkpoint_size")
      import :: c_ptr, c_int, c_size_t
      type(c_ptr), value :: h
      integer(c_size_t), intent(out) :: nbytes
    end function
    integer(c_int) function dyc_checkpoint_save(h, buf, nbytes, loc) bind(c, name="gfsdc_checkpoint_save")
      import :: c_ptr, c_int, c_size_t
      type(c_ptr), value :: h, buf
      integer(c_size_t), value :: nbytes
      integer(c_int), value :: loc
    end function
    integer(c_int) function dyc_checkpoint_load(h, buf, nbytes, loc) bind(c, name="gfsdc_checkpoint_load")
      import :: c_ptr, c_int, c_size_t
      type(c_ptr), value :: h, buf
      integer(c_size_t), value :: nbytes
      integer(c_int), value :: loc
    end function
    ! stoch_data.f90 init_stochdata
    subroutine dyc_stoch_defaults(s) bind(c, name="gfsdc_stoch_defaults")
      import :: gfsdc_stoch_cfg
      type(gfsdc_stoch_cfg), intent(inout) :: s
    end subroutine
    integer(c_int) function dyc_stoch_init(h, s) bind(c, name="gfsdc_stoch_init")
      import :: c_ptr, c_int, gfsdc_stoch_cfg
      type(c_ptr), value :: h
      type(gfsdc_stoch_cfg), intent(in) :: s
    end function
    integer(c_int) function dyc_get_stoch(h, which, out, dtype, loc) bind(c, name="gfsdc_get_stoch")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, out
      integer(c_int), value :: which, dtype, loc
    end function
    ! AR(1) temperature pattern of every member (ndimspec complex per member)
    integer(c_int) function dyc_get_pattern(h, spec, dtype, loc) bind(c, name="gfsdc_get_pattern")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, spec
      integer(c_int), value :: dtype, loc
    end function
    ! learned column physics in place of the simple physics' column call
    integer(c_int) function dyc_mlp_init(h, hidden, nlayers, maxbatch, capacity, seed) &
        bind(c, name="gfsdc_mlp_init")
      import :: c_ptr, c_int, c_long, c_long_long
      type(c_ptr), value :: h
      integer(c_int), value :: hidden, nlayers, maxbatch
      integer(c_long), value :: capacity
      integer(c_long_long), value :: seed
    end function
    integer(c_int) function dyc_mlp_collect(h, nsteps, stride, ncollected) bind(c, name="gfsdc_mlp_collect")
      import :: c_ptr, c_int, c_long
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps, stride
      integer(c_long) :: ncollected
    end function
    integer(c_int) function dyc_mlp_normalize(h) bind(c, name="gfsdc_mlp_normalize")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
    end function
    integer(c_int) function dyc_mlp_train(h, nsteps, batch, lr, wd, ntrain, losses) bind(c, name="gfsdc_mlp_train")
      import :: c_ptr, c_int, c_long, c_double
      type(c_ptr), value :: h, losses
      integer(c_int), value :: nsteps, batch
      real(c_double), value :: lr, wd
      integer(c_long), value :: ntrain
    end function
    integer(c_int) function dyc_mlp_enable(h, on, judge_stride, threshold) bind(c, name="gfsdc_mlp_enable")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: on, judge_stride
      real(c_double), value :: threshold
    end function
    integer(c_int) function dyc_mlp_judge_stats(h, out3) bind(c, name="gfsdc_mlp_judge_stats")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, out3
    end function
    ! parameters and normalisation of the network (get and / or set; c_null_ptr
    ! skips a direction): a network trained in one context run in another
    integer(c_int) function dyc_mlp_params(h, get, set) bind(c, name="gfsdc_mlp_params")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, get, set
    end function
    integer(c_int) function dyc_mlp_norm(h, get, set) bind(c, name="gfsdc_mlp_norm")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, get, set
    end function
  end interface

contains

  ! stop unless this bind(c) gfsdc_cfg has the library's size: a type that
  ! misses fields would let gfsdc_cfg_defaults write past it
  subroutine dyc_check_abi()
    type(gfsdc_cfg) :: probe
    if (c_sizeof(probe) /= dyc_cfg_size()) then
      print *, 'gfsdc: gfsdc_cfg is', c_sizeof(probe), 'bytes here and', dyc_cfg_size(), &
               'in libgfsdc (include/gfsdc.h); rebuild the Fortran layer against the header'
      stop 1
    end if
  end subroutine
end module gfsdc_dycore
