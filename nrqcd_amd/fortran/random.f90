!> random module (random.f90:126-253) backed by the library's ranlxd.
module random
  use, intrinsic :: iso_fortran_env, only: int64, real64
  use nrq_capi
  implicit none
  private
  public :: InitModule, IsModuleInitialised, GetRandomUniformReal
  character(len=6), parameter, public :: modulename = 'random'
  logical :: initialised = .false.
contains
  !> random::InitModule(CustomSeed) -> ranlxd_init(2, seed)
  subroutine InitModule(CustomSeed)
    integer(int64), intent(in), optional :: CustomSeed
    integer(int64) :: seed
    seed = 1
    if (present(CustomSeed)) seed = CustomSeed
    call nrq_check(nrq_random_init(seed), 'random::InitModule')
    initialised = .true.
  end subroutine InitModule

  logical function IsModuleInitialised()
    IsModuleInitialised = initialised
  end function IsModuleInitialised

  function GetRandomUniformReal(n) result(res)
    integer(int64), intent(in) :: n
    real(real64) :: res(n)
    call nrq_check(nrq_random_uniform(res, n), 'GetRandomUniformReal')
  end function GetRandomUniformReal
end module random
