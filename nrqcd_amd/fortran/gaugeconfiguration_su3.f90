!> gaugeconfiguration_su3 (gaugeconfiguration_su3.f90:33-116) with the
!! configuration resident in GPU memory.  The type keeps the reference's name
!! and type-bound procedures; its only component is the library handle.
!! Intrinsic assignment of the reference (a deep copy, main.f90:2320) is a
!! defined assignment here, and a final procedure releases the device memory.
module gaugeconfiguration_su3
  use, intrinsic :: iso_fortran_env, only: int8, int64
  use, intrinsic :: iso_c_binding
  use precision, only: fp
  use nrq_capi
  implicit none
  private
  public :: GaugeConfiguration, UpdateEnsemble

  character(len=22), parameter, public :: modulename = 'gaugeconfiguration_su3'
  integer(int8), parameter, public :: nSUN = 3_int8, nGen = 8_int8

  type GaugeConfiguration
     type(c_ptr), private :: handle = c_null_ptr
     integer(c_int), private :: storage = NRQ_PRECISION_FP64
   contains
     procedure, public :: Allocate
     procedure, public :: Deallocate
     procedure, public :: SetStoragePrecision
     procedure, public :: ColdInit
     procedure, public :: TransversePolarisedOccupiedInit_Box
     procedure, public :: GetLink_G
     procedure, public :: GetLink_M
     procedure, public :: GetEfield_G
     procedure, public :: GetDeviationFromGausslaw
     procedure, public :: GetGaussLawViolation
     procedure, public :: GetEnergy
     procedure, public :: GetMeanPlaquette
     procedure, public :: GetPlaquette_G
     procedure, public :: GetPlaquette_M
     procedure, public :: GetElectricField_AlgebraCoordinate
     generic, public :: GetElectricField => GetElectricField_AlgebraCoordinate
     procedure, public :: Upload
     procedure, public :: Download
     procedure, public :: Reunitarise
     procedure, public :: SetReunitarisationInterval
     procedure, public :: CoulombGaugefixing
     procedure, public :: EquilibriumInit
     procedure, public :: TransversePolarisedOccupiedInit
     procedure, public :: TransversePolarisedOccupiedInit_specificSender
     procedure, public :: GetTransverseAACorrelator
     procedure, public :: GetTransverseEECorrelator
     procedure, private :: Update_Leapfrog
     generic, public :: Update => Update_Leapfrog
     procedure, private :: Assign
     generic, public :: assignment(=) => Assign
     procedure, public :: CHandle
     final :: Finalise
  end type GaugeConfiguration

contains

  !> GaugeConfiguration%Allocate, gaugeconfiguration_su3.f90:128-138
  subroutine Allocate(GaugeConf)
    class(GaugeConfiguration), intent(inout) :: GaugeConf
    if (c_associated(GaugeConf%handle)) return
    call nrq_check(nrq_gauge_create(GaugeConf%storage, GaugeConf%handle), 'GaugeConfiguration%Allocate')
  end subroutine Allocate

  !> GaugeConfiguration%Deallocate, gaugeconfiguration_su3.f90:145-151
  subroutine Deallocate(GaugeConf)
    class(GaugeConfiguration), intent(inout) :: GaugeConf
    integer(c_int) :: rc
    if (c_associated(GaugeConf%handle)) rc = nrq_gauge_destroy(GaugeConf%handle)
    GaugeConf%handle = c_null_ptr
  end subroutine Deallocate

  subroutine Finalise(GaugeConf)
    type(GaugeConfiguration), intent(inout) :: GaugeConf
    integer(c_int) :: rc
    if (c_associated(GaugeConf%handle)) rc = nrq_gauge_destroy(GaugeConf%handle)
    GaugeConf%handle = c_null_ptr
  end subroutine Finalise

  !> Selects FP64 (default, the reference's) or FP32 storage of links and E
  !! for the next allocation.
  subroutine SetStoragePrecision(GaugeConf, bits)
    class(GaugeConfiguration), intent(inout) :: GaugeConf
    integer, intent(in) :: bits
    call GaugeConf%Deallocate
    GaugeConf%storage = int(bits, c_int)
  end subroutine SetStoragePrecision

  function CHandle(GaugeConf)
    class(GaugeConfiguration), intent(in) :: GaugeConf
    type(c_ptr) :: CHandle
    CHandle = GaugeConf%handle
  end function CHandle

  !> Deep copy (the reference's intrinsic assignment, main.f90:2320)
  subroutine Assign(to, from)
    class(GaugeConfiguration), intent(inout) :: to
    class(GaugeConfiguration), intent(in) :: from
    to%storage = from%storage
    if (.not. c_associated(from%handle)) then
       call to%Deallocate
       return
    end if
    call to%Allocate
    call nrq_check(nrq_gauge_copy(to%handle, from%handle), 'GaugeConfiguration assignment')
  end subroutine Assign

  !> ColdInit, gaugeconfiguration_su3.f90:363-378
  subroutine ColdInit(GaugeConf)
    class(GaugeConfiguration), intent(inout) :: GaugeConf
    call GaugeConf%Allocate
    call nrq_check(nrq_gauge_cold(GaugeConf%handle), 'ColdInit')
  end subroutine ColdInit

  !> TransversePolarisedOccupiedInit_Box(SaturationScale, Amplitude, Coupling,
  !! [aa_correlator], [ee_correlator]), gaugeconfiguration_su3.f90:906-951
  !! (draws from the random module's stream); the optional outputs are the
  !! transverse correlators of the momentum-space fields, one per momentum index.
  subroutine TransversePolarisedOccupiedInit_Box(GaugeConf, SaturationScale, Amplitude, Coupling, &
       aa_correlator, ee_correlator)
    use lattice, only: GetLocalLatticeSize
    class(GaugeConfiguration), intent(inout) :: GaugeConf
    real(fp), intent(in) :: SaturationScale, Amplitude, Coupling
    real(fp), optional, allocatable, target, intent(out) :: aa_correlator(:), ee_correlator(:)
    type(c_ptr) :: pa, pe
    call GaugeConf%Allocate
    pa = c_null_ptr
    pe = c_null_ptr
    if (present(aa_correlator)) then
       allocate(aa_correlator(GetLocalLatticeSize()))
       pa = c_loc(aa_correlator)
    end if
    if (present(ee_correlator)) then
       allocate(ee_correlator(GetLocalLatticeSize()))
       pe = c_loc(ee_correlator)
    end if
    call nrq_check(nrq_gauge_box_ic_correlators(GaugeConf%handle, SaturationScale, Amplitude, Coupling, pa, pe), &
         'TransversePolarisedOccupiedInit_Box')
  end subroutine TransversePolarisedOccupiedInit_Box

  !> EquilibriumInit(Beta, nefieldinit, nequilibrium, [tolerance_Eprojection],
  !! [ChargeDensity], [MeasureEnergy], [filename]),
  !! gaugeconfiguration_su3.f90:380-814: thermal E projected onto the Gauss
  !! law and leapfrog equilibration, nefieldinit times from cold links.  As in
  !! the reference, the Gauss deviation after each iteration is printed and,
  !! with MeasureEnergy, the iteration number and energy are written to
  !! filename and printed.
  subroutine EquilibriumInit(GaugeConf, Beta, nefieldinit, nequilibrium, tolerance_Eprojection, ChargeDensity, &
       MeasureEnergy, filename)
    use, intrinsic :: iso_fortran_env, only: output_unit
    use mpiinterface, only: MPIstop
    class(GaugeConfiguration), intent(inout) :: GaugeConf
    real(fp), intent(in) :: Beta
    integer(int64), intent(in) :: nefieldinit, nequilibrium
    real(fp), intent(in), optional :: tolerance_Eprojection
    real(fp), intent(in), optional, target, contiguous :: ChargeDensity(:, :)
    logical, intent(in), optional :: MeasureEnergy
    character(len=*), intent(in), optional :: filename
    real(fp), allocatable, target :: energies(:), deviations(:)
    integer(c_int64_t), allocatable, target :: projections(:)
    real(fp) :: tol
    type(c_ptr) :: prho
    logical :: measure
    integer :: fid
    integer(int64) :: it
    tol = -1
    if (present(tolerance_Eprojection)) tol = tolerance_Eprojection
    prho = c_null_ptr
    if (present(ChargeDensity)) prho = c_loc(ChargeDensity)
    measure = .false.
    if (present(MeasureEnergy)) measure = MeasureEnergy
    if (measure .and. .not. present(filename)) call MPIstop('Filename for energy output missing')
    allocate(energies(max(nefieldinit, 1_int64)), deviations(max(nefieldinit, 1_int64)), &
         projections(max(nefieldinit, 1_int64)))
    call GaugeConf%Allocate
    call nrq_check(nrq_gauge_equilibrium_init(GaugeConf%handle, Beta, nefieldinit, nequilibrium, tol, prho, &
         c_loc(energies), c_loc(deviations), c_loc(projections)), 'EquilibriumInit')
    if (measure) open(newunit=fid, file=trim(filename), status='replace', form='formatted', action='write')
    do it = 1, nefieldinit
       if (measure) then
          write(fid, *) it, energies(it)
          write(output_unit, *) it, energies(it)
       end if
       write(output_unit, *) 'G[U,E]=', deviations(it)
    end do
    if (measure) close(fid)
  end subroutine EquilibriumInit

  !> TransversePolarisedOccupiedInit(Occupation, [aa_correlator],
  !! [ee_correlator]), gaugeconfiguration_su3.f90:1263-1446: any occupation
  !! per lattice index (private in the reference, behind the box form)
  subroutine TransversePolarisedOccupiedInit(GaugeConf, Occupation, aa_correlator, ee_correlator)
    class(GaugeConfiguration), intent(inout) :: GaugeConf
    real(fp), intent(in), target, contiguous :: Occupation(:)
    real(fp), optional, allocatable, target, intent(out) :: aa_correlator(:), ee_correlator(:)
    call transverse_init(GaugeConf, Occupation, 0_c_int, 'TransversePolarisedOccupiedInit', aa_correlator, ee_correlator)
  end subroutine TransversePolarisedOccupiedInit

  !> TransversePolarisedOccupiedInit_specificSender(Occupation,
  !! [aa_correlator], [ee_correlator]), gaugeconfiguration_su3.f90:959-1255:
  !! amplitudes drawn in momentum space in lattice order, independent of the
  !! number of processes, then made Hermitian (private in the reference)
  subroutine TransversePolarisedOccupiedInit_specificSender(GaugeConf, Occupation, aa_correlator, ee_correlator)
    class(GaugeConfiguration), intent(inout) :: GaugeConf
    real(fp), intent(in), target, contiguous :: Occupation(:)
    real(fp), optional, allocatable, target, intent(out) :: aa_correlator(:), ee_correlator(:)
    call transverse_init(GaugeConf, Occupation, 1_c_int, 'TransversePolarisedOccupiedInit_specificSender', &
         aa_correlator, ee_correlator)
  end subroutine TransversePolarisedOccupiedInit_specificSender

  subroutine transverse_init(GaugeConf, Occupation, sender, what, aa_correlator, ee_correlator)
    use lattice, only: GetLocalLatticeSize
    class(GaugeConfiguration), intent(inout) :: GaugeConf
    real(fp), intent(in), target, contiguous :: Occupation(:)
    integer(c_int), intent(in) :: sender
    character(len=*), intent(in) :: what
    real(fp), optional, allocatable, target, intent(out) :: aa_correlator(:), ee_correlator(:)
    type(c_ptr) :: pa, pe
    if (size(Occupation, kind=int64) /= GetLocalLatticeSize()) then
       write(*, '(A,A)') 'Error in ', what
       write(*, '(A,I0,A,I0)') 'occupation has ', size(Occupation, kind=int64), ' entries, the lattice ', &
            GetLocalLatticeSize()
       error stop 1
    end if
    call GaugeConf%Allocate
    pa = c_null_ptr
    pe = c_null_ptr
    if (present(aa_correlator)) then
       allocate(aa_correlator(GetLocalLatticeSize()))
       pa = c_loc(aa_correlator)
    end if
    if (present(ee_correlator)) then
       allocate(ee_correlator(GetLocalLatticeSize()))
       pe = c_loc(ee_correlator)
    end if
    call nrq_check(nrq_gauge_transverse_init(GaugeConf%handle, c_loc(Occupation), sender, pa, pe), what)
  end subroutine transverse_init

  !> GetLink_G(i, LatticeIndex), gaugeconfiguration_su3.f90:160-196
  function GetLink_G(GaugeConf, i, LatticeIndex)
    class(GaugeConfiguration), intent(in) :: GaugeConf
    integer(int8), intent(in) :: i
    integer(int64), intent(in) :: LatticeIndex
    complex(fp) :: GetLink_G(nSUN, nSUN)
    call nrq_check(nrq_gauge_get_link(GaugeConf%handle, int(i, c_int), LatticeIndex, GetLink_G), 'GetLink_G')
  end function GetLink_G

  function GetLink_M(GaugeConf, i, MemoryIndex)
    class(GaugeConfiguration), intent(in) :: GaugeConf
    integer(int8), intent(in) :: i
    integer(int64), intent(in) :: MemoryIndex
    complex(fp) :: GetLink_M(nSUN, nSUN)
    GetLink_M = GaugeConf%GetLink_G(i, MemoryIndex)
  end function GetLink_M

  !> E^a_i in lattice units (GetEfield_G, gaugeconfiguration_su3.f90:239-254)
  function GetEfield_G(GaugeConf, i, LatticeIndex)
    class(GaugeConfiguration), intent(in) :: GaugeConf
    integer(int8), intent(in) :: i
    integer(int64), intent(in) :: LatticeIndex
    real(fp) :: GetEfield_G(nGen)
    call nrq_check(nrq_gauge_get_efield(GaugeConf%handle, int(i, c_int), LatticeIndex, GetEfield_G), 'GetEfield_G')
  end function GetEfield_G

  !> GetElectricField_AlgebraCoordinate(a, i, LatticeIndex),
  !! gaugeconfiguration_su3.f90:2028-2049: E^a_i / (a_i a_0)
  real(fp) function GetElectricField_AlgebraCoordinate(GaugeConf, a, i, LatticeIndex)
    use lattice, only: GetLatticeSpacing
    class(GaugeConfiguration), intent(in) :: GaugeConf
    integer(int8), intent(in) :: a, i
    integer(int64), intent(in) :: LatticeIndex
    real(fp) :: e(nGen)
    e = GaugeConf%GetEfield_G(i, LatticeIndex)
    GetElectricField_AlgebraCoordinate = e(a) / GetLatticeSpacing(i) / GetLatticeSpacing(0_int8)
  end function GetElectricField_AlgebraCoordinate

  function observables(GaugeConf)
    class(GaugeConfiguration), intent(in) :: GaugeConf
    real(fp) :: observables(4)
    call nrq_check(nrq_gauge_observables(GaugeConf%handle, observables), 'gauge observables')
  end function observables

  !> GetDeviationFromGaussLaw(), gaugeconfiguration_su3.f90:1480-1531
  real(fp) function GetDeviationFromGausslaw(GaugeConf)
    class(GaugeConfiguration), intent(in) :: GaugeConf
    real(fp) :: o(4)
    o = observables(GaugeConf)
    GetDeviationFromGausslaw = o(2)
  end function GetDeviationFromGausslaw

  !> Gauss-law constraint max_{x,a} 2|Tr(T^a G(x))| (GetGdeviation,
  !! gaugeconfiguration_su3.f90:724-813)
  real(fp) function GetGaussLawViolation(GaugeConf)
    class(GaugeConfiguration), intent(in) :: GaugeConf
    real(fp) :: o(4)
    o = observables(GaugeConf)
    GetGaussLawViolation = o(3)
  end function GetGaussLawViolation

  !> GetEnergy(), gaugeconfiguration_su3.f90:1769-1817
  real(fp) function GetEnergy(GaugeConf)
    class(GaugeConfiguration), intent(in) :: GaugeConf
    real(fp) :: o(4)
    o = observables(GaugeConf)
    GetEnergy = o(1)
  end function GetEnergy

  real(fp) function GetMeanPlaquette(GaugeConf)
    class(GaugeConfiguration), intent(in) :: GaugeConf
    real(fp) :: o(4)
    o = observables(GaugeConf)
    GetMeanPlaquette = o(4)
  end function GetMeanPlaquette

  !> GetPlaquette_G(i, j, LatticeIndex) for spatial i, j,
  !! gaugeconfiguration_su3.f90:1591-1615
  function GetPlaquette_G(GaugeConf, i, j, LatticeIndex)
    class(GaugeConfiguration), intent(in) :: GaugeConf
    integer(int8), intent(in) :: i, j
    integer(int64), intent(in) :: LatticeIndex
    complex(fp) :: GetPlaquette_G(nSUN, nSUN)
    call nrq_check(nrq_gauge_plaquette(GaugeConf%handle, int(i, c_int), int(j, c_int), LatticeIndex, &
         GetPlaquette_G), 'GetPlaquette_G')
  end function GetPlaquette_G

  function GetPlaquette_M(GaugeConf, i, j, MemoryIndex)
    class(GaugeConfiguration), intent(in) :: GaugeConf
    integer(int8), intent(in) :: i, j
    integer(int64), intent(in) :: MemoryIndex
    complex(fp) :: GetPlaquette_M(nSUN, nSUN)
    GetPlaquette_M = GaugeConf%GetPlaquette_G(i, j, MemoryIndex)
  end function GetPlaquette_M

  !> Update([StepWidth]) -> Update_Leapfrog, gaugeconfiguration_su3.f90:2082-2111;
  !! nsteps (default 1) repeats it on the device without host round trips.
  subroutine Update_Leapfrog(GaugeConf, StepWidth_, nsteps)
    class(GaugeConfiguration), intent(inout) :: GaugeConf
    real(fp), intent(in), optional :: StepWidth_
    integer(int64), intent(in), optional :: nsteps
    real(fp) :: sw
    integer(int64) :: n
    sw = 1._fp
    if (present(StepWidth_)) sw = StepWidth_
    n = 1
    if (present(nsteps)) n = nsteps
    call nrq_check(nrq_gauge_leapfrog(GaugeConf%handle, sw, n), 'GaugeConfiguration%Update')
  end subroutine Update_Leapfrog

  !> Update for an ensemble of configurations of one storage precision: one
  !! launch per pass over all members, replayed from a captured hipGraph.
  subroutine UpdateEnsemble(GaugeConfs, StepWidth_, nsteps)
    type(GaugeConfiguration), intent(inout) :: GaugeConfs(:)
    real(fp), intent(in), optional :: StepWidth_
    integer(int64), intent(in), optional :: nsteps
    type(c_ptr), allocatable :: h(:)
    real(fp) :: sw
    integer(int64) :: n
    integer :: k
    sw = 1._fp
    if (present(StepWidth_)) sw = StepWidth_
    n = 1
    if (present(nsteps)) n = nsteps
    allocate(h(size(GaugeConfs)))
    do k = 1, size(GaugeConfs)
       h(k) = GaugeConfs(k)%handle
    end do
    call nrq_check(nrq_gauge_leapfrog_batch(h, int(size(GaugeConfs), c_int), sw, n), 'UpdateEnsemble')
  end subroutine UpdateEnsemble

  !> Copies links(3,3,3,V) / efield(8,3,V) in the reference layout to the device.
  subroutine Upload(GaugeConf, links, efield)
    class(GaugeConfiguration), intent(inout) :: GaugeConf
    complex(fp), intent(in), target :: links(:, :, :, :)
    real(fp), intent(in), target :: efield(:, :, :)
    call GaugeConf%Allocate
    call nrq_check(nrq_gauge_upload(GaugeConf%handle, c_loc(links), c_loc(efield)), 'GaugeConfiguration%Upload')
  end subroutine Upload

  subroutine Download(GaugeConf, links, efield)
    class(GaugeConfiguration), intent(in) :: GaugeConf
    complex(fp), intent(out), target :: links(:, :, :, :)
    real(fp), intent(out), target :: efield(:, :, :)
    call nrq_check(nrq_gauge_download(GaugeConf%handle, c_loc(links), c_loc(efield)), 'GaugeConfiguration%Download')
  end subroutine Download

  subroutine Reunitarise(GaugeConf)
    class(GaugeConfiguration), intent(inout) :: GaugeConf
    call nrq_check(nrq_gauge_reunitarise(GaugeConf%handle), 'GaugeConfiguration%Reunitarise')
  end subroutine Reunitarise

  !> Reunitarise after every Interval leapfrog steps (0: never).
  subroutine SetReunitarisationInterval(GaugeConf, Interval)
    class(GaugeConfiguration), intent(inout) :: GaugeConf
    integer(int64), intent(in) :: Interval
    call GaugeConf%Allocate
    call nrq_check(nrq_gauge_set_reunitarise_interval(GaugeConf%handle, Interval), &
         'GaugeConfiguration%SetReunitarisationInterval')
  end subroutine SetReunitarisationInterval

  !> GaugeConfiguration%CoulombGaugefixing(Tolerance_, alpha_, MaxIterations_),
  !! gaugeconfiguration_su3.f90:2269-2311, with the reference's defaults, and
  !! CoulombGaugefixing_Links :2318-2497 (Davies' Fourier-accelerated steepest
  !! descent on the links; E is not transformed).  Prints the final loop index
  !! as the reference does (transformations applied + 1).
  subroutine CoulombGaugefixing(GaugeConf, Tolerance_, alpha_, MaxIterations_)
    use, intrinsic :: iso_fortran_env, only: output_unit
    class(GaugeConfiguration), intent(inout) :: GaugeConf
    real(fp), intent(in), optional :: Tolerance_
    real(fp), intent(in), optional :: alpha_
    integer(int64), intent(in), optional :: MaxIterations_
    real(fp) :: Tolerance, alpha, deviation
    integer(int64) :: MaxIterations, done
    if (present(Tolerance_)) then
       Tolerance = Tolerance_
    else
       Tolerance = 1.E-8
    end if
    if (present(alpha_)) then
       alpha = alpha_
    else
       alpha = 0.08_fp
    end if
    if (present(MaxIterations_)) then
       MaxIterations = MaxIterations_
    else
       MaxIterations = 100
    end if
    call nrq_check(nrq_gauge_coulomb_gaugefix(GaugeConf%handle, Tolerance, alpha, MaxIterations, done, deviation), &
         'GaugeConfiguration%CoulombGaugefixing')
    write(output_unit, *) deviation
    write(output_unit, *) done + 1, 'iterations in Coulomb gauge fixing performed'
    call flush(output_unit)
  end subroutine CoulombGaugefixing

  !> GaugeConfiguration%GetTransverseAACorrelator(Correlator),
  !! gaugeconfiguration_su3.f90:1825-1856: one value per momentum index
  subroutine GetTransverseAACorrelator(GaugeConf, Correlator)
    use lattice, only: GetLocalLatticeSize
    class(GaugeConfiguration), intent(in) :: GaugeConf
    real(fp), allocatable, target, intent(out) :: Correlator(:)
    allocate(Correlator(GetLocalLatticeSize()))
    call nrq_check(nrq_gauge_transverse_correlators(GaugeConf%handle, c_loc(Correlator), c_null_ptr), &
         'GaugeConfiguration%GetTransverseAACorrelator')
  end subroutine GetTransverseAACorrelator

  !> GaugeConfiguration%GetTransverseEECorrelator(Correlator),
  !! gaugeconfiguration_su3.f90:1863-1894: one value per momentum index
  subroutine GetTransverseEECorrelator(GaugeConf, Correlator)
    use lattice, only: GetLocalLatticeSize
    class(GaugeConfiguration), intent(in) :: GaugeConf
    real(fp), allocatable, target, intent(out) :: Correlator(:)
    allocate(Correlator(GetLocalLatticeSize()))
    call nrq_check(nrq_gauge_transverse_correlators(GaugeConf%handle, c_null_ptr, c_loc(Correlator)), &
         'GaugeConfiguration%GetTransverseEECorrelator')
  end subroutine GetTransverseEECorrelator
end module gaugeconfiguration_su3
