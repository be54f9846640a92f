!> Working precision of the host-side Fortran (the reference fixes real64,
!! precision.f90:16); the device storage precision is chosen per
!! GaugeConfiguration.
module precision
  use, intrinsic :: iso_fortran_env, only: real64
  implicit none
  public
  integer, parameter :: fp = real64
end module precision
