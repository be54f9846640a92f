!> lattice module (lattice.f90) on one GPU: index, neighbour and momentum
!! maps come from the library (periodic, 1-based, x fastest); memory index
!! and lattice index coincide because there is no partition and no halo.
module lattice
  use, intrinsic :: iso_fortran_env, only: int8, int64
  use, intrinsic :: iso_c_binding
  use precision, only: fp
  use nrq_capi
  implicit none
  private
  public :: InitModule, FinalizeModule, IsModuleInitialised, GetLatticeIndex, GetLatticePosition, &
       GetNeib_G, GetNeib_M, GetLatticeIndex_M, GetMemoryIndex, GetNegativeLatticeIndex, &
       GetMomentum_G, GetMomentum_M, GetNorm2Momentum_G, GetNorm2Momentum_M, GetMaxNorm2Momentum, GetLatticeSize, &
       GetMemorySize, GetLocalLatticeSize, GetLatticeExtension, GetLatticeSpacing, GetVolume, &
       GetProc_G, GetProc_M

  integer(int8), parameter, public :: nDim = 3_int8
  character(len=7), parameter, public :: modulename = 'lattice'

  logical :: initialised = .false.
  integer(int64) :: extensions(3) = -1
  real(fp) :: spacings(0:3) = -1

contains
  !> lattice::InitModule(LatticeExtensions_, LatticeSpacings_), lattice.f90:124-228
  subroutine InitModule(LatticeExtensions_, LatticeSpacings_)
    integer(int64), intent(in) :: LatticeExtensions_(3)
    real(fp), intent(in) :: LatticeSpacings_(0:3)
    call nrq_check(nrq_lattice_init(LatticeExtensions_, LatticeSpacings_), 'lattice::InitModule')
    extensions = LatticeExtensions_
    spacings = LatticeSpacings_
    initialised = .true.
  end subroutine InitModule

  subroutine FinalizeModule()
    call nrq_check(nrq_lattice_finalize(), 'lattice::FinalizeModule')
    initialised = .false.
  end subroutine FinalizeModule

  logical function IsModuleInitialised()
    IsModuleInitialised = initialised
  end function IsModuleInitialised

  integer(int64) function GetLatticeIndex(Position)
    integer(int64), intent(in) :: Position(3)
    call nrq_check(nrq_lattice_index(Position, GetLatticeIndex), 'GetLatticeIndex')
  end function GetLatticeIndex

  function GetLatticePosition(LatticeIndex)
    integer(int64), intent(in) :: LatticeIndex
    integer(int64) :: GetLatticePosition(3)
    call nrq_check(nrq_lattice_position(LatticeIndex, GetLatticePosition), 'GetLatticePosition')
  end function GetLatticePosition

  integer(int64) function GetNeib_G(i, LatticeIndex)
    integer(int8), intent(in) :: i
    integer(int64), intent(in) :: LatticeIndex
    call nrq_check(nrq_lattice_neib(int(i, c_int), LatticeIndex, GetNeib_G), 'GetNeib_G')
  end function GetNeib_G

  integer(int64) function GetNeib_M(i, MemoryIndex)
    integer(int8), intent(in) :: i
    integer(int64), intent(in) :: MemoryIndex
    GetNeib_M = GetNeib_G(i, MemoryIndex)
  end function GetNeib_M

  pure integer(int64) function GetLatticeIndex_M(MemoryIndex)
    integer(int64), intent(in) :: MemoryIndex
    GetLatticeIndex_M = MemoryIndex
  end function GetLatticeIndex_M

  pure integer(int64) function GetMemoryIndex(LatticeIndex)
    integer(int64), intent(in) :: LatticeIndex
    GetMemoryIndex = LatticeIndex
  end function GetMemoryIndex

  integer(int64) function GetNegativeLatticeIndex(LatticeIndex)
    integer(int64), intent(in) :: LatticeIndex
    call nrq_check(nrq_lattice_negative_index(LatticeIndex, GetNegativeLatticeIndex), 'GetNegativeLatticeIndex')
  end function GetNegativeLatticeIndex

  function GetMomentum_G(LatticeIndex)
    integer(int64), intent(in) :: LatticeIndex
    complex(fp) :: GetMomentum_G(3)
    call nrq_check(nrq_lattice_momentum(LatticeIndex, GetMomentum_G), 'GetMomentum_G')
  end function GetMomentum_G

  function GetMomentum_M(MemoryIndex)
    integer(int64), intent(in) :: MemoryIndex
    complex(fp) :: GetMomentum_M(3)
    GetMomentum_M = GetMomentum_G(MemoryIndex)
  end function GetMomentum_M

  real(fp) function GetNorm2Momentum_G(LatticeIndex)
    integer(int64), intent(in) :: LatticeIndex
    GetNorm2Momentum_G = norm2(abs(GetMomentum_G(LatticeIndex)))
  end function GetNorm2Momentum_G

  !> lattice::GetMaxNorm2Momentum, lattice.f90:1101-1105
  real(fp) function GetMaxNorm2Momentum()
    call nrq_check(nrq_lattice_max_norm2_momentum(GetMaxNorm2Momentum), 'lattice::GetMaxNorm2Momentum')
  end function GetMaxNorm2Momentum

  real(fp) function GetNorm2Momentum_M(MemoryIndex)
    integer(int64), intent(in) :: MemoryIndex
    GetNorm2Momentum_M = GetNorm2Momentum_G(MemoryIndex)
  end function GetNorm2Momentum_M

  pure integer(int64) function GetLatticeSize()
    GetLatticeSize = product(extensions)
  end function GetLatticeSize

  pure integer(int64) function GetMemorySize()
    GetMemorySize = product(extensions)
  end function GetMemorySize

  pure integer(int64) function GetLocalLatticeSize()
    GetLocalLatticeSize = product(extensions)
  end function GetLocalLatticeSize

  pure integer(int64) function GetLatticeExtension(i)
    integer(int8), intent(in) :: i
    GetLatticeExtension = extensions(i)
  end function GetLatticeExtension

  pure real(fp) function GetLatticeSpacing(i)
    integer(int8), intent(in) :: i
    GetLatticeSpacing = spacings(i)
  end function GetLatticeSpacing

  pure real(fp) function GetVolume()
    GetVolume = product(extensions * spacings(1:3))
  end function GetVolume

  pure integer function GetProc_G(LatticeIndex)
    integer(int64), intent(in) :: LatticeIndex
    GetProc_G = 0
  end function GetProc_G

  pure integer function GetProc_M(MemoryIndex)
    integer(int64), intent(in) :: MemoryIndex
    GetProc_M = 0
  end function GetProc_M
end module lattice
