! Real kinds for building the Fortran host layer standalone. When the layer is
! linked into the reference model, the reference's own src/kinds.f90 is used
! instead (same names: r_single, r_double, r_kind, default_real).
module kinds
  use iso_c_binding, only: c_float, c_double
  implicit none
  integer, parameter :: r_single = c_float, r_double = c_double
  integer, parameter :: default_real = 1
  integer, parameter :: r_kind = r_single
end module kinds
