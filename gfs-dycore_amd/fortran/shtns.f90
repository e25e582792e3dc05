! module shtns: drop-in replacement of the reference's transform wrapper
! (src/shtns.f90) on top of libgfsdc (include/gfsdc.h), written from scratch.
!
! Same public names, argument meaning and module data as src/shtns.f90:58-60:
!   shtns_init(nlon,nlat,ntrunc[,nthreads,polar_opt])        (shtns.f90:90)
!   grdtospec(datagrid,dataspec)                            (shtns.f90:171)
!   spectogrd(dataspec,datagrid)                            (shtns.f90:203)
!   getuv(vrtspec,divspec,ugrid,vgrid,rsphere)              (shtns.f90:233)
!   getvrtdivspec(ugrid,vgrid,vrtspec,divspec,rsphere)      (shtns.f90:275)
!   getgrad(dataspec,gradx,grady,rsphere)                   (shtns.f90:313)
!   shtns_destroy()                                         (shtns.f90:158)
!   lats, lons, gauwts, areawts, nlm, degree, order, lap, invlap,
!   current_nlon, current_nlat, current_ntrunc
! Each transform also accepts a batch: dataspec(ndimspec,nfield) with
! datagrid(nlon,nlat,nfield), which replaces an OpenMP loop over levels by
! one GPU launch sequence. ntrunc is inferred from size(dataspec,1) and a
! change of grid or truncation re-initialises, as in the reference
! (shtns.f90:183-188). Errors stop the program, as the reference does.
! The context computes in FP64 (as SHTns) unless GFSDC_PRECISION=fp32.
module shtns
  use iso_c_binding
  use kinds, only: r_kind, r_double
  ! the one bind(c) definition of gfsdc_cfg, checked against the library's size
  use gfsdc_dycore, only: gfsdc_cfg, dyc_defaults, dyc_create, dyc_destroy, dyc_check_abi
  implicit none
  private
  public :: shtns_init, grdtospec, spectogrd, getuv, getvrtdivspec, getgrad, shtns_destroy
  public :: gauwts, lats, lons, nlm, degree, order, lap, invlap, &
            current_nlon, current_nlat, current_ntrunc, areawts

  integer, save :: current_nlon = -1, current_nlat = -1, current_ntrunc = -1
  integer :: nlm
  real(r_kind), dimension(:), allocatable :: lap, invlap, gauwts
  real(r_kind), dimension(:,:), allocatable :: lats, lons, areawts
  integer, dimension(:), allocatable :: degree, order
  type(c_ptr), save :: ctx = c_null_ptr

  integer(c_int), parameter :: GFSDC_HOST = 0, DT_F64 = 0, DT_F32 = 1
  integer(c_int), parameter :: GFSDC_FP64 = 0, GFSDC_FP32 = 1

  interface
    integer(c_int) function c_griddata(h, lats, gauwts, areawts, degree, order, lap, invlap) &
        bind(c, name="gfsdc_griddata")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, lats, gauwts, areawts, degree, order, lap, invlap
    end function
    integer(c_int) function c_nlm(h) bind(c, name="gfsdc_nlm")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
    end function
    integer(c_int) function c_spectogrd(h, spec, grid, nf, dtype, loc) bind(c, name="gfsdc_spectogrd")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, spec, grid
      integer(c_int), value :: nf, dtype, loc
    end function
    integer(c_int) function c_grdtospec(h, grid, spec, nf, dtype, loc) bind(c, name="gfsdc_grdtospec")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, grid, spec
      integer(c_int), value :: nf, dtype, loc
    end function
    integer(c_int) function c_getuv(h, vrt, div, u, v, nf, rs, dtype, loc) bind(c, name="gfsdc_getuv")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h, vrt, div, u, v
      integer(c_int), value :: nf, dtype, loc
      real(c_double), value :: rs
    end function
    integer(c_int) function c_getvrtdivspec(h, u, v, vrt, div, nf, rs, dtype, loc) &
        bind(c, name="gfsdc_getvrtdivspec")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h, u, v, vrt, div
      integer(c_int), value :: nf, dtype, loc
      real(c_double), value :: rs
    end function
    integer(c_int) function c_getgrad(h, spec, gx, gy, nf, rs, dtype, loc) bind(c, name="gfsdc_getgrad")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h, spec, gx, gy
      integer(c_int), value :: nf, dtype, loc
      real(c_double), value :: rs
    end function
    type(c_ptr) function c_last_error() bind(c, name="gfsdc_last_error")
      import :: c_ptr
    end function
  end interface

  interface grdtospec
    module procedure grdtospec_1, grdtospec_n
  end interface
  interface spectogrd
    module procedure spectogrd_1, spectogrd_n
  end interface
  interface getuv
    module procedure getuv_1, getuv_n
  end interface
  interface getvrtdivspec
    module procedure getvrtdivspec_1, getvrtdivspec_n
  end interface
  interface getgrad
    module procedure getgrad_1, getgrad_n
  end interface

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(c_last_error(), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    print *, 'gfsdc error in ', what, ': ', msg(1:n)
    stop 1
  end subroutine

  integer(c_int) function dtype_of()
    if (r_kind == c_float) then
      dtype_of = DT_F32
    else
      dtype_of = DT_F64
    end if
  end function

  integer function trunc_of(nspec)
    integer, intent(in) :: nspec
    trunc_of = nint((-1. + sqrt(1 + 8 * real(nspec))) / 2.) - 1   ! shtns.f90:183
  end function

  subroutine shtns_init(nlon, nlat, ntrunc, nthreads, polar_opt)
    integer, intent(in) :: nlon, nlat, ntrunc
    integer, intent(in), optional :: nthreads
    real(r_double), intent(in), optional :: polar_opt
    type(gfsdc_cfg) :: cfg
    character(len=16) :: prec
    integer :: i, stat
    real(c_double), allocatable, target :: lat1(:), w1(:), aw(:,:), lp(:), ilp(:)
    integer(c_int), allocatable, target :: dg(:), od(:)
    real(r_double) :: pi
    call shtns_destroy()
    call dyc_check_abi()
    call dyc_defaults(cfg)
    cfg%nlon = nlon
    cfg%nlat = nlat
    cfg%ntrunc = ntrunc
    cfg%nlev = 0
    cfg%precision = GFSDC_FP64
    call get_environment_variable("GFSDC_PRECISION", prec, status=stat)
    if (stat == 0 .and. (prec == "fp32" .or. prec == "FP32")) cfg%precision = GFSDC_FP32
    call check(dyc_create(cfg, ctx), "shtns_init")
    nlm = c_nlm(ctx)
    if (nlm /= (ntrunc + 1) * (ntrunc + 2) / 2) then
      print *, 'error: nlm not what expected', nlm, ntrunc
      stop
    end if
    current_nlon = nlon
    current_nlat = nlat
    current_ntrunc = ntrunc
    allocate(lat1(nlat), w1(nlat), aw(nlon, nlat), lp(nlm), ilp(nlm), dg(nlm), od(nlm))
    call check(c_griddata(ctx, c_loc(lat1), c_loc(w1), c_loc(aw), c_loc(dg), c_loc(od), &
                          c_loc(lp), c_loc(ilp)), "griddata")
    allocate(lats(nlon, nlat), lons(nlon, nlat), gauwts(nlat), areawts(nlon, nlat))
    allocate(degree(nlm), order(nlm), lap(nlm), invlap(nlm))
    pi = 4.d0 * atan(1.d0)
    do i = 1, nlon
      lats(i, :) = lat1
      lons(i, :) = 2.d0 * pi * real(i - 1, r_double) / nlon
    end do
    gauwts = w1
    areawts = aw
    degree = dg
    order = od
    lap = lp
    invlap = ilp
    if (present(nthreads)) continue
    if (present(polar_opt)) continue
  end subroutine

  subroutine shtns_destroy()
    integer(c_int) :: rc
    if (c_associated(ctx)) rc = dyc_destroy(ctx)
    ctx = c_null_ptr
    current_nlon = -1
    current_nlat = -1
    current_ntrunc = -1
    if (allocated(lats)) deallocate(lats)
    if (allocated(lons)) deallocate(lons)
    if (allocated(gauwts)) deallocate(gauwts)
    if (allocated(areawts)) deallocate(areawts)
    if (allocated(degree)) deallocate(degree)
    if (allocated(order)) deallocate(order)
    if (allocated(lap)) deallocate(lap)
    if (allocated(invlap)) deallocate(invlap)
  end subroutine

  subroutine ensure(nlon, nlat, nspec)
    integer, intent(in) :: nlon, nlat, nspec
    integer :: ntrunc
    ntrunc = trunc_of(nspec)
    if (nlon /= current_nlon .or. nlat /= current_nlat .or. ntrunc /= current_ntrunc) &
      call shtns_init(nlon, nlat, ntrunc)
  end subroutine

  ! ---------------------------------------------------------------- batched
  subroutine grdtospec_n(datagrid, dataspec)
    real(r_kind), dimension(:,:,:), intent(in), contiguous, target :: datagrid
    complex(r_kind), dimension(:,:), intent(out), contiguous, target :: dataspec
    call ensure(size(datagrid, 1), size(datagrid, 2), size(dataspec, 1))
    !$omp critical (gfsdc)
    call check(c_grdtospec(ctx, c_loc(datagrid), c_loc(dataspec), int(size(datagrid, 3), c_int), &
                           dtype_of(), GFSDC_HOST), "grdtospec")
    !$omp end critical (gfsdc)
  end subroutine

  subroutine spectogrd_n(dataspec, datagrid)
    complex(r_kind), dimension(:,:), intent(in), contiguous, target :: dataspec
    real(r_kind), dimension(:,:,:), intent(out), contiguous, target :: datagrid
    call ensure(size(datagrid, 1), size(datagrid, 2), size(dataspec, 1))
    !$omp critical (gfsdc)
    call check(c_spectogrd(ctx, c_loc(dataspec), c_loc(datagrid), int(size(datagrid, 3), c_int), &
                           dtype_of(), GFSDC_HOST), "spectogrd")
    !$omp end critical (gfsdc)
  end subroutine

  subroutine getuv_n(vrtspec, divspec, ugrid, vgrid, rsphere)
    real(r_kind), intent(in) :: rsphere
    complex(r_kind), dimension(:,:), intent(in), contiguous, target :: vrtspec, divspec
    real(r_kind), dimension(:,:,:), intent(out), contiguous, target :: ugrid, vgrid
    call ensure(size(ugrid, 1), size(ugrid, 2), size(vrtspec, 1))
    !$omp critical (gfsdc)
    call check(c_getuv(ctx, c_loc(vrtspec), c_loc(divspec), c_loc(ugrid), c_loc(vgrid), &
                       int(size(ugrid, 3), c_int), real(rsphere, c_double), dtype_of(), GFSDC_HOST), "getuv")
    !$omp end critical (gfsdc)
  end subroutine

  subroutine getvrtdivspec_n(ugrid, vgrid, vrtspec, divspec, rsphere)
    real(r_kind), intent(in) :: rsphere
    real(r_kind), dimension(:,:,:), intent(in), contiguous, target :: ugrid, vgrid
    complex(r_kind), dimension(:,:), intent(out), contiguous, target :: vrtspec, divspec
    call ensure(size(ugrid, 1), size(ugrid, 2), size(vrtspec, 1))
    !$omp critical (gfsdc)
    call check(c_getvrtdivspec(ctx, c_loc(ugrid), c_loc(vgrid), c_loc(vrtspec), c_loc(divspec), &
                               int(size(ugrid, 3), c_int), real(rsphere, c_double), dtype_of(), GFSDC_HOST), &
               "getvrtdivspec")
    !$omp end critical (gfsdc)
  end subroutine

  subroutine getgrad_n(dataspec, ugrid, vgrid, rsphere)
    real(r_kind), intent(in) :: rsphere
    complex(r_kind), dimension(:,:), intent(in), contiguous, target :: dataspec
    real(r_kind), dimension(:,:,:), intent(out), contiguous, target :: ugrid, vgrid
    call ensure(size(ugrid, 1), size(ugrid, 2), size(dataspec, 1))
    !$omp critical (gfsdc)
    call check(c_getgrad(ctx, c_loc(dataspec), c_loc(ugrid), c_loc(vgrid), int(size(ugrid, 3), c_int), &
                         real(rsphere, c_double), dtype_of(), GFSDC_HOST), "getgrad")
    !$omp end critical (gfsdc)
  end subroutine

  ! ------------------------------------------------- one field (reference API)
  subroutine grdtospec_1(datagrid, dataspec)
    real(r_kind), dimension(:,:), intent(in), contiguous, target :: datagrid
    complex(r_kind), dimension(:), intent(out), contiguous, target :: dataspec
    real(r_kind), pointer :: g3(:,:,:)
    complex(r_kind), pointer :: s2(:,:)
    call c_f_pointer(c_loc(datagrid), g3, [size(datagrid, 1), size(datagrid, 2), 1])
    call c_f_pointer(c_loc(dataspec), s2, [size(dataspec), 1])
    call grdtospec_n(g3, s2)
  end subroutine

  subroutine spectogrd_1(dataspec, datagrid)
    complex(r_kind), dimension(:), intent(in), contiguous, target :: dataspec
    real(r_kind), dimension(:,:), intent(out), contiguous, target :: datagrid
    real(r_kind), pointer :: g3(:,:,:)
    complex(r_kind), pointer :: s2(:,:)
    call c_f_pointer(c_loc(datagrid), g3, [size(datagrid, 1), size(datagrid, 2), 1])
    call c_f_pointer(c_loc(dataspec), s2, [size(dataspec), 1])
    call spectogrd_n(s2, g3)
  end subroutine

  subroutine getuv_1(vrtspec, divspec, ugrid, vgrid, rsphere)
    real(r_kind), intent(in) :: rsphere
    complex(r_kind), dimension(:), intent(in), contiguous, target :: vrtspec, divspec
    real(r_kind), dimension(:,:), intent(out), contiguous, target :: ugrid, vgrid
    real(r_kind), pointer :: u3(:,:,:), v3(:,:,:)
    complex(r_kind), pointer :: a2(:,:), b2(:,:)
    call c_f_pointer(c_loc(ugrid), u3, [size(ugrid, 1), size(ugrid, 2), 1])
    call c_f_pointer(c_loc(vgrid), v3, [size(vgrid, 1), size(vgrid, 2), 1])
    call c_f_pointer(c_loc(vrtspec), a2, [size(vrtspec), 1])
    call c_f_pointer(c_loc(divspec), b2, [size(divspec), 1])
    call getuv_n(a2, b2, u3, v3, rsphere)
  end subroutine

  subroutine getvrtdivspec_1(ugrid, vgrid, vrtspec, divspec, rsphere)
    real(r_kind), intent(in) :: rsphere
    real(r_kind), dimension(:,:), intent(in), contiguous, target :: ugrid, vgrid
    complex(r_kind), dimension(:), intent(out), contiguous, target :: vrtspec, divspec
    real(r_kind), pointer :: u3(:,:,:), v3(:,:,:)
    complex(r_kind), pointer :: a2(:,:), b2(:,:)
    call c_f_pointer(c_loc(ugrid), u3, [size(ugrid, 1), size(ugrid, 2), 1])
    call c_f_pointer(c_loc(vgrid), v3, [size(vgrid, 1), size(vgrid, 2), 1])
    call c_f_pointer(c_loc(vrtspec), a2, [size(vrtspec), 1])
    call c_f_pointer(c_loc(divspec), b2, [size(divspec), 1])
    call getvrtdivspec_n(u3, v3, a2, b2, rsphere)
  end subroutine

  subroutine getgrad_1(dataspec, ugrid, vgrid, rsphere)
    real(r_kind), intent(in) :: rsphere
    complex(r_kind), dimension(:), intent(in), contiguous, target :: dataspec
    real(r_kind), dimension(:,:), intent(out), contiguous, target :: ugrid, vgrid
    real(r_kind), pointer :: u3(:,:,:), v3(:,:,:)
    complex(r_kind), pointer :: a2(:,:)
    call c_f_pointer(c_loc(ugrid), u3, [size(ugrid, 1), size(ugrid, 2), 1])
    call c_f_pointer(c_loc(vgrid), v3, [size(vgrid, 1), size(vgrid, 2), 1])
    call c_f_pointer(c_loc(dataspec), a2, [size(dataspec), 1])
    call getgrad_n(a2, u3, v3, rsphere)
  end subroutine

end module shtns
