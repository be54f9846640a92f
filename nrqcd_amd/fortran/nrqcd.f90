!> nrqcd module (nrqcd.f90:14-117) with the propagators resident in GPU
!! memory and the Crank-Nicolson solve done by the library's Krylov solver
!! (replaces PETSc KSP).  QuarkProp/AntiQprop are reached through
!! GetQuarkProp_G/GetAntiQProp_G and Download instead of public components.
module nrqcd
  use, intrinsic :: iso_fortran_env, only: int8, int64
  use, intrinsic :: iso_c_binding
  use precision, only: fp
  use nrq_capi
  use gaugeconfiguration_su3, only: SU3GaugeConfiguration => GaugeConfiguration
  implicit none
  private
  public :: NRQCDField, InitModule, FinalizeModule, UpdateBatch

  character(len=5), parameter, public :: modulename = 'nrqcd'
  integer(int8), parameter, public :: nDof = 6_int8
  integer(int8), parameter, public :: nWilsonCoefficients = 4_int8

  type NRQCDField
     type(c_ptr), private :: handle = c_null_ptr
   contains
     procedure, public :: Destructor
     procedure, public :: InitSinglePoint
     procedure, public :: InitSinglePointSingleDoF
     procedure, public :: GetNorm_Quark
     procedure, public :: GetNorm_AntiQ
     procedure, private :: Update_CrankNicholson
     generic, public :: Update => Update_CrankNicholson
     procedure, public :: GetMesoncorrelator_O1_1S0_ZeroMomentum
     procedure, public :: GetMesoncorrelator_O1_3S1_ZeroMomentum
     procedure, public :: GetMesoncorrelator_O8_1S0_ZeroMomentum
     procedure, public :: GetMesoncorrelator_O8_3S1_ZeroMomentum
     procedure, public :: GetMesoncorrelator_O1_1S0_ZeroMomentum_gconjg
     procedure, public :: GetMesoncorrelator_O1_3S1_ZeroMomentum_gconjg
     procedure, public :: GetMesoncorrelator_O8_1S0_ZeroMomentum_gconjg
     procedure, public :: GetMesoncorrelator_O8_3S1_ZeroMomentum_gconjg
     procedure, public :: GetQuarkProp_G
     procedure, public :: GetAntiQProp_G
     procedure, public :: Download
     procedure, public :: Upload
     procedure, private :: Assign
     generic, public :: assignment(=) => Assign
     final :: Finalise
  end type NRQCDField

  logical :: initialised = .false.
  integer(int64), public :: TotalKrylovIterations = 0

contains
  !> nrqcd::InitModule(Mass, WilsonCoeffs, StepWidth, kspTol), nrqcd.f90:167-192
  subroutine InitModule(Mass, WilsonCoeffs, StepWidth, kspTol)
    real(fp), intent(in) :: Mass, StepWidth, kspTol
    complex(fp), intent(in) :: WilsonCoeffs(nWilsonCoefficients)
    call nrq_check(nrq_cn_init(Mass, WilsonCoeffs, StepWidth, kspTol), 'nrqcd::InitModule')
    initialised = .true.
  end subroutine InitModule

  !> nrqcd::FinalizeModule, nrqcd.f90:199-212
  subroutine FinalizeModule()
    call nrq_check(nrq_cn_finalize(), 'nrqcd::FinalizeModule')
    initialised = .false.
  end subroutine FinalizeModule

  subroutine ensure(HeavyField)
    class(NRQCDField), intent(inout) :: HeavyField
    if (.not. c_associated(HeavyField%handle)) &
         call nrq_check(nrq_heavy_create(HeavyField%handle), 'NRQCDField allocation')
  end subroutine ensure

  subroutine Destructor(object)
    class(NRQCDField), intent(inout) :: object
    integer(c_int) :: rc
    if (c_associated(object%handle)) rc = nrq_heavy_destroy(object%handle)
    object%handle = c_null_ptr
  end subroutine Destructor

  subroutine Finalise(object)
    type(NRQCDField), intent(inout) :: object
    integer(c_int) :: rc
    if (c_associated(object%handle)) rc = nrq_heavy_destroy(object%handle)
    object%handle = c_null_ptr
  end subroutine Finalise

  !> assignment(=) -> assign, nrqcd.f90:219-228 (deep copy)
  subroutine Assign(to, from)
    class(NRQCDField), intent(inout) :: to
    class(NRQCDField), intent(in) :: from
    call ensure(to)
    call nrq_check(nrq_heavy_copy(to%handle, from%handle), 'NRQCDField assignment')
  end subroutine Assign

  !> InitSinglePoint(LatticeIndex_quark, LatticeIndex_antiq), nrqcd.f90:324-350
  subroutine InitSinglePoint(HeavyField, LatticeIndex_quark, LatticeIndex_antiq)
    class(NRQCDField), intent(inout) :: HeavyField
    integer(int64), intent(in) :: LatticeIndex_quark, LatticeIndex_antiq
    call ensure(HeavyField)
    call nrq_check(nrq_heavy_init_point(HeavyField%handle, LatticeIndex_quark, LatticeIndex_antiq), &
         'InitSinglePoint')
  end subroutine InitSinglePoint

  !> InitSinglePointSingleDoF, nrqcd.f90:357-393
  subroutine InitSinglePointSingleDoF(HeavyField, spin_quark, colour_quark, LatticeIndex_quark, &
       spin_antiq, colour_antiq, LatticeIndex_antiq)
    class(NRQCDField), intent(inout) :: HeavyField
    integer(int8), intent(in) :: spin_quark, spin_antiq, colour_quark, colour_antiq
    integer(int64), intent(in) :: LatticeIndex_quark, LatticeIndex_antiq
    call ensure(HeavyField)
    call nrq_check(nrq_heavy_init_point_dof(HeavyField%handle, int(spin_quark, c_int), int(colour_quark, c_int), &
         LatticeIndex_quark, int(spin_antiq, c_int), int(colour_antiq, c_int), LatticeIndex_antiq), &
         'InitSinglePointSingleDoF')
  end subroutine InitSinglePointSingleDoF

  !> Update(GaugeConf, Mass, WilsonCoeffs, [StepWidth]) -> Update_CrankNicholson,
  !! nrqcd.f90:1053-1168
  subroutine Update_CrankNicholson(HeavyField, GaugeConf, Mass, WilsonCoeffs, StepWidth)
    class(NRQCDField), intent(inout) :: HeavyField
    type(SU3GaugeConfiguration), intent(in) :: GaugeConf
    real(fp), intent(in) :: Mass
    complex(fp), intent(in) :: WilsonCoeffs(nWilsonCoefficients)
    real(fp), intent(in), optional :: StepWidth
    real(fp) :: sw
    type(nrq_solve_stats) :: st
    sw = 1._fp
    if (present(StepWidth)) sw = StepWidth
    call nrq_check(nrq_heavy_update(HeavyField%handle, GaugeConf%CHandle(), Mass, WilsonCoeffs, sw, st), &
         'NRQCDField%Update')
    TotalKrylovIterations = TotalKrylovIterations + st%iterations
  end subroutine Update_CrankNicholson

  !> Update for several heavy fields with their own masses on one gauge
  !! configuration, as one batched multi-right-hand-side solve.
  subroutine UpdateBatch(HeavyFields, GaugeConf, Masses, WilsonCoeffs, StepWidth)
    type(NRQCDField), intent(inout) :: HeavyFields(:)
    type(SU3GaugeConfiguration), intent(in) :: GaugeConf
    real(fp), intent(in) :: Masses(:)
    complex(fp), intent(in) :: WilsonCoeffs(nWilsonCoefficients)
    real(fp), intent(in), optional :: StepWidth
    type(c_ptr), allocatable :: h(:)
    real(fp) :: sw
    type(nrq_solve_stats) :: st
    integer :: k
    sw = 1._fp
    if (present(StepWidth)) sw = StepWidth
    if (size(Masses) /= size(HeavyFields)) error stop 'UpdateBatch: one mass per field'
    allocate(h(size(HeavyFields)))
    do k = 1, size(HeavyFields)
       h(k) = HeavyFields(k)%handle
    end do
    call nrq_check(nrq_heavy_update_batch(h, Masses, int(size(HeavyFields), c_int), GaugeConf%CHandle(), &
         WilsonCoeffs, sw, st), 'UpdateBatch')
    TotalKrylovIterations = TotalKrylovIterations + st%iterations
  end subroutine UpdateBatch

  !> GetNorm_Quark, nrqcd.f90:547-557
  real(fp) function GetNorm_Quark(HeavyField)
    class(NRQCDField), intent(in) :: HeavyField
    real(fp) :: na
    call nrq_check(nrq_heavy_norms(HeavyField%handle, GetNorm_Quark, na), 'GetNorm_Quark')
  end function GetNorm_Quark

  !> GetNorm_AntiQ, nrqcd.f90:564-572
  real(fp) function GetNorm_AntiQ(HeavyField)
    class(NRQCDField), intent(in) :: HeavyField
    real(fp) :: nq
    call nrq_check(nrq_heavy_norms(HeavyField%handle, nq, GetNorm_AntiQ), 'GetNorm_AntiQ')
  end function GetNorm_AntiQ

  function correlators(HeavyField)
    class(NRQCDField), intent(in) :: HeavyField
    complex(fp) :: correlators(4)
    call nrq_check(nrq_heavy_correlators(HeavyField%handle, correlators), 'meson correlators')
  end function correlators

  !> nrqcd.f90:1484-1517
  complex(fp) function GetMesoncorrelator_O1_1S0_ZeroMomentum(HeavyField) result(res)
    class(NRQCDField), intent(in) :: HeavyField
    complex(fp) :: c(4)
    c = correlators(HeavyField)
    res = c(1)
  end function GetMesoncorrelator_O1_1S0_ZeroMomentum

  !> nrqcd.f90:1525-1566
  complex(fp) function GetMesoncorrelator_O1_3S1_ZeroMomentum(HeavyField) result(res)
    class(NRQCDField), intent(in) :: HeavyField
    complex(fp) :: c(4)
    c = correlators(HeavyField)
    res = c(2)
  end function GetMesoncorrelator_O1_3S1_ZeroMomentum

  !> nrqcd.f90:1574-1611
  complex(fp) function GetMesoncorrelator_O8_1S0_ZeroMomentum(HeavyField) result(res)
    class(NRQCDField), intent(in) :: HeavyField
    complex(fp) :: c(4)
    c = correlators(HeavyField)
    res = c(3)
  end function GetMesoncorrelator_O8_1S0_ZeroMomentum

  !> nrqcd.f90:1619-1661
  complex(fp) function GetMesoncorrelator_O8_3S1_ZeroMomentum(HeavyField) result(res)
    class(NRQCDField), intent(in) :: HeavyField
    complex(fp) :: c(4)
    c = correlators(HeavyField)
    res = c(4)
  end function GetMesoncorrelator_O8_3S1_ZeroMomentum

  function correlators_gconjg(HeavyField)
    class(NRQCDField), intent(in) :: HeavyField
    complex(fp) :: correlators_gconjg(4)
    call nrq_check(nrq_heavy_correlators_gconjg(HeavyField%handle, correlators_gconjg), 'meson correlators (gconjg)')
  end function correlators_gconjg

  !> nrqcd.f90:1728-1761
  complex(fp) function GetMesoncorrelator_O1_1S0_ZeroMomentum_gconjg(HeavyField) result(res)
    class(NRQCDField), intent(in) :: HeavyField
    complex(fp) :: c(4)
    c = correlators_gconjg(HeavyField)
    res = c(1)
  end function GetMesoncorrelator_O1_1S0_ZeroMomentum_gconjg

  !> nrqcd.f90:1769-1809
  complex(fp) function GetMesoncorrelator_O1_3S1_ZeroMomentum_gconjg(HeavyField) result(res)
    class(NRQCDField), intent(in) :: HeavyField
    complex(fp) :: c(4)
    c = correlators_gconjg(HeavyField)
    res = c(2)
  end function GetMesoncorrelator_O1_3S1_ZeroMomentum_gconjg

  !> nrqcd.f90:1817-1857
  complex(fp) function GetMesoncorrelator_O8_1S0_ZeroMomentum_gconjg(HeavyField) result(res)
    class(NRQCDField), intent(in) :: HeavyField
    complex(fp) :: c(4)
    c = correlators_gconjg(HeavyField)
    res = c(3)
  end function GetMesoncorrelator_O8_1S0_ZeroMomentum_gconjg

  !> nrqcd.f90:1865-1905
  complex(fp) function GetMesoncorrelator_O8_3S1_ZeroMomentum_gconjg(HeavyField) result(res)
    class(NRQCDField), intent(in) :: HeavyField
    complex(fp) :: c(4)
    c = correlators_gconjg(HeavyField)
    res = c(4)
  end function GetMesoncorrelator_O8_3S1_ZeroMomentum_gconjg

  !> Copies both propagators (6,6,V) to the host.
  subroutine Download(HeavyField, QuarkProp, AntiQProp)
    class(NRQCDField), intent(in) :: HeavyField
    complex(fp), intent(out), target :: QuarkProp(:, :, :), AntiQProp(:, :, :)
    call nrq_check(nrq_heavy_download(HeavyField%handle, c_loc(QuarkProp), c_loc(AntiQProp)), 'NRQCDField%Download')
  end subroutine Download

  subroutine Upload(HeavyField, QuarkProp, AntiQProp)
    class(NRQCDField), intent(inout) :: HeavyField
    complex(fp), intent(in), target :: QuarkProp(:, :, :), AntiQProp(:, :, :)
    call ensure(HeavyField)
    call nrq_check(nrq_heavy_upload(HeavyField%handle, c_loc(QuarkProp), c_loc(AntiQProp)), 'NRQCDField%Upload')
  end subroutine Upload

  !> GetQuarkProp_G(LatticeIndex), nrqcd.f90:399-413 (downloads the whole
  !! field; meant for occasional access, not for loops over sites)
  function GetQuarkProp_G(HeavyField, LatticeIndex)
    use lattice, only: GetLatticeSize
    class(NRQCDField), intent(in) :: HeavyField
    integer(int64), intent(in) :: LatticeIndex
    complex(fp) :: GetQuarkProp_G(nDof, nDof)
    complex(fp), allocatable :: q(:, :, :), a(:, :, :)
    allocate(q(nDof, nDof, GetLatticeSize()), a(nDof, nDof, GetLatticeSize()))
    call HeavyField%Download(q, a)
    GetQuarkProp_G = q(:, :, LatticeIndex)
  end function GetQuarkProp_G

  !> GetAntiQProp_G(LatticeIndex), nrqcd.f90:420-434
  function GetAntiQProp_G(HeavyField, LatticeIndex)
    use lattice, only: GetLatticeSize
    class(NRQCDField), intent(in) :: HeavyField
    integer(int64), intent(in) :: LatticeIndex
    complex(fp) :: GetAntiQProp_G(nDof, nDof)
    complex(fp), allocatable :: q(:, :, :), a(:, :, :)
    allocate(q(nDof, nDof, GetLatticeSize()), a(nDof, nDof, GetLatticeSize()))
    call HeavyField%Download(q, a)
    GetAntiQProp_G = a(:, :, LatticeIndex)
  end function GetAntiQProp_G
end module nrqcd
