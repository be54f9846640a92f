!> nrqcd_run -- Fortran driver over the GPU library, with the reference's
!! positional command line (main.f90:3301-3340) for the hot-path modes:
!!   1  MeasureGluondistribution           (main.f90:3020-3190)
!!        nrqcd_run 1 Nx Ny Nz a0 ax ay az CoMTime Seed Qs Amp g
!!                  GaugefixingMaxIterations GaugefixingTolerance GaugefixingCoefficient FileGluondist
!!        -> FileGluondist: |p| and sqrt(AA EE) per momentum '(2(SP,E13.6,1X))'
!!   2  MeasureEnergyAndGaussLawDeviation   (main.f90:2572-2718)
!!        nrqcd_run 2 Nx Ny Nz a0 ax ay az CoMTime TimeRange Seed Qs Amp g
!!        -> energy_gauss.txt: time, energy, Gauss deviation '(3(SP,E13.6,1X))'
!!   3  MeasureHeavyQuarkoniumCorrelators  (main.f90:2169-2447)
!!        nrqcd_run 3 Nx Ny Nz ds ax ay az tmin tmax smax Seed Qs Amp g kspTol M
!!                  c1re c1im c2re c2im c3re c3im c4re c4im FileCorrelator FileNorm
!!   33 MeasureHeavyQuarkoniumCorrelators_Equilibrium (main.f90:18-280)
!!        nrqcd_run 33 Nx Ny Nz a0 ax ay az smax Seed Beta nefieldinit nequilibrium kspTol M
!!                  c1re c1im c2re c2im c3re c3im c4re c4im FileCorrelator FileNorm [MeasureEnergy [FileEnergy]]
!!        -> as mode 32, from EquilibriumInit instead of cold links
!!   4  ComputeSpectrum (main.f90:1865-2100)
!!        nrqcd_run 4 nHeaderLines FileInput FileOutput MaxTime
!!        -> FileOutput: omega and 2 Im of the transformed windowed correlator
!!           '(2(SP,E13.6,1X))', negative frequencies first; input.txt,
!!           symm.txt, hanned.txt and fftformat.txt as in the reference
!!   31 MeasureHeavyQuarkoniumCorrelators_oneT (main.f90:524-830)
!!        nrqcd_run 31 Nx Ny Nz a0 ax ay az t smax Seed Qs Amp g kspTol M
!!                  c1re c1im c2re c2im c3re c3im c4re c4im FileCorrelator FileNorm FileCorrelator_gconjg
!!        -> all four correlators (and their _gconjg variants) against s at fixed t
!!   32 MeasureHeavyQuarkoniumCorrelators_free (main.f90:282-522)
!!        nrqcd_run 32 Nx Ny Nz a0 ax ay az smax Seed kspTol M
!!                  c1re c1im c2re c2im c3re c3im c4re c4im FileCorrelator FileNorm
!!        -> all four correlators against s in a cold gauge field
!! Numbers are read list-directed.  The reference reads fixed formats
!! ('(I4)', '(F10.13)', '(E15.7)', '(F6.1)'); both give the same values for
!! the inputs of its templates (reals written with a decimal point, at most
!! 10 characters), and list-directed input also takes reals such as 1 or
!! 1e-12 at face value.  An optional trailing argument 'fp32' in mode 2
!! selects FP32 link/E storage.
program nrqcd_run
  use, intrinsic :: iso_fortran_env, only: int8, int64, output_unit
  use, intrinsic :: ieee_arithmetic
  use precision, only: fp
  use mpiinterface, only: InitModule_MPI => InitModule, ThisProc, MPIstop
  use lattice, only: InitModule_Lattice => InitModule, GetLatticeSpacing, GetNorm2Momentum_G
  use random, only: InitModule_Random => InitModule
  use gaugeconfiguration_su3, only: GaugeConfiguration
  use nrqcd, only: NRQCDField, InitModule_NRQCD => InitModule, FinalizeModule_NRQCD => FinalizeModule, &
       TotalKrylovIterations
  use twfft, only: GetSpectralFunction
  implicit none

  character(len=256) :: arg
  integer :: mode

  if (command_argument_count() < 1) call usage()
  call get_command_argument(1, arg)
  read(arg, *) mode
  call InitModule_MPI()
  select case (mode)
  case (1)
     call MeasureGluondistribution()
  case (2)
     call MeasureEnergyAndGaussLawDeviation()
  case (3)
     call MeasureHeavyQuarkoniumCorrelators()
  case (4)
     call ComputeSpectrum()
  case (31)
     call MeasureHeavyQuarkoniumCorrelators_oneT()
  case (32)
     call MeasureHeavyQuarkoniumCorrelators_free()
  case (33)
     call MeasureHeavyQuarkoniumCorrelators_Equilibrium()
  case default
     call usage()
  end select

contains

  subroutine usage()
    write(output_unit, '(A)') 'usage: nrqcd_run 1|2|3|31|32|33|4 <positional parameters of the reference mode>'
    stop 2
  end subroutine usage

  integer(int64) function iarg(k)
    integer, intent(in) :: k
    call get_command_argument(k, arg)
    read(arg, *) iarg
  end function iarg

  real(fp) function rarg(k)
    integer, intent(in) :: k
    call get_command_argument(k, arg)
    read(arg, *) rarg
  end function rarg

  !> Mode 1: gluon distribution of the box initial condition evolved to
  !! CoMTime: Coulomb gauge fixing of a copy, transverse AA and EE
  !! correlators, and per momentum |p| and sqrt(AA EE)
  subroutine MeasureGluondistribution()
    integer(int64) :: ext(3), seed, TimeSteps, GFMaxIterations, is
    real(fp) :: spacings(0:3), CoMTime, qs, amp, g, GFTolerance, GFCoefficient
    character(len=256) :: fname
    type(GaugeConfiguration) :: GaugeConf, GaugeConf_gaugefixed
    real(fp), allocatable :: aa(:), ee(:)
    integer :: k, fid

    if (command_argument_count() < 17) call usage()
    do k = 1, 3
       ext(k) = iarg(1 + k)
    end do
    do k = 0, 3
       spacings(k) = rarg(5 + k)
    end do
    CoMTime = rarg(9)
    seed = iarg(10)
    qs = rarg(11)
    amp = rarg(12)
    g = rarg(13)
    GFMaxIterations = iarg(14)
    GFTolerance = rarg(15)
    GFCoefficient = rarg(16)
    call get_command_argument(17, fname)
    TimeSteps = ceiling(abs(CoMTime) / spacings(0))

    call InitModule_Lattice(ext, spacings)
    call InitModule_Random(seed + ThisProc())
    call GaugeConf%TransversePolarisedOccupiedInit_Box(qs, amp, g)
    if (TimeSteps > 0) call GaugeConf%Update(sign(1._fp, CoMTime), TimeSteps)

    GaugeConf_gaugefixed = GaugeConf
    call GaugeConf_gaugefixed%CoulombGaugefixing(Tolerance_=GFTolerance, alpha_=GFCoefficient, &
         MaxIterations_=GFMaxIterations)
    call GaugeConf_gaugefixed%GetTransverseAACorrelator(aa)
    call GaugeConf_gaugefixed%GetTransverseEECorrelator(ee)

    open(newunit=fid, file=trim(fname), status='replace', form='formatted', action='write')
    do is = 1, size(aa)
       write(fid, '(2(SP,E13.6,1X))') GetNorm2Momentum_G(is), sqrt(aa(is) * ee(is))
    end do
    close(fid)
    call GaugeConf_gaugefixed%Deallocate
    call GaugeConf%Deallocate
    write(output_unit, *) 'Simulation completed'
  end subroutine MeasureGluondistribution

  !> Mode 2: energy and Gauss-law deviation along the evolution
  subroutine MeasureEnergyAndGaussLawDeviation()
    integer(int64) :: ext(3), seed, TimeSteps, it
    real(fp) :: spacings(0:3), CoMTime, TimeRange, qs, amp, g, time, energy, gauss, gmax, plaq
    type(GaugeConfiguration) :: GaugeConf
    integer :: k, fid, fid2

    if (command_argument_count() < 14) call usage()
    do k = 1, 3
       ext(k) = iarg(1 + k)
    end do
    do k = 0, 3
       spacings(k) = rarg(5 + k)
    end do
    CoMTime = rarg(9)
    TimeRange = rarg(10)
    seed = iarg(11)
    qs = rarg(12)
    amp = rarg(13)
    g = rarg(14)
    TimeSteps = ceiling(TimeRange / spacings(0))
    if (command_argument_count() >= 15) then
       call get_command_argument(15, arg)
       if (trim(arg) == 'fp32') call GaugeConf%SetStoragePrecision(32)
    end if

    call InitModule_Lattice(ext, spacings)
    call InitModule_Random(seed + ThisProc())
    call GaugeConf%TransversePolarisedOccupiedInit_Box(qs, amp, g)

    open(newunit=fid, file='energy_gauss.txt', status='replace', form='formatted', action='write')
    open(newunit=fid2, file='energy_gauss_constraint_plaquette.txt', status='replace', form='formatted', &
         action='write')
    do it = 0, TimeSteps
       time = it * GetLatticeSpacing(0_int8)
       gauss = GaugeConf%GetDeviationFromGaussLaw()
       energy = GaugeConf%GetEnergy()
       gmax = GaugeConf%GetGaussLawViolation()
       plaq = GaugeConf%GetMeanPlaquette()
       write(fid, '(3(SP,E13.6,1X))') time, energy, gauss
       write(fid2, '(5(SP,E23.16,1X))') time, energy, gauss, gmax, plaq
       write(output_unit, *) time
       call GaugeConf%Update()
    end do
    close(fid)
    close(fid2)
    call GaugeConf%Deallocate
    write(output_unit, *) 'Simulation completed'
  end subroutine MeasureEnergyAndGaussLawDeviation

  !> Mode 4: spectral function of the correlator in columns 1-3 (t, Re, Im)
  !! of a file (main.f90:1865-2100).  Rows are read after nHeaderLines header
  !! lines until t > MaxTime or the end of the file, dt is the difference of
  !! the first two times (ReadSignal, :2018-2062).  The symmetrised signal
  !! (negative times -conj(C)), its Hann window, the reordering and the
  !! backward transform with scale dt run on the GPU; the spectrum is
  !! 2 Im of the result, written with its frequencies (:2066-2099).
  subroutine ComputeSpectrum()
    character(len=256) :: fin, fout
    integer :: nheader, k, fid, stat
    integer(int64) :: nsig, i, nn
    real(fp) :: maxtime, t, t1, re, im, dt, wmin, dw
    complex(fp), allocatable :: Correlator(:), symm(:), hanned(:), fftformat(:)
    real(fp), allocatable :: Spectrum(:)
    real(fp), parameter :: pi = acos(-1._fp)

    if (command_argument_count() < 5) call usage()
    nheader = int(iarg(2))
    call get_command_argument(3, fin)
    call get_command_argument(4, fout)
    maxtime = rarg(5)
    write(output_unit, *) '**********************'
    write(output_unit, *) 'Extraction of spectrum'
    write(output_unit, *) '**********************'
    write(output_unit, *)
    write(output_unit, *) 'Input file:  ', trim(fin)
    write(output_unit, *) 'Output file: ', trim(fout)

    open(newunit=fid, file=trim(fin), status='old', form='formatted', action='read')
    do k = 1, nheader
       read(fid, *)
    end do
    nsig = 0
    t = -huge(t)
    do
       read(fid, *, iostat=stat) t, re, im
       if (t > maxtime) exit
       if (stat /= 0) exit
       nsig = nsig + 1
    end do
    close(fid)
    if (nsig < 2) call MPIstop('ComputeSpectrum: the signal needs at least two time points')
    allocate(Correlator(nsig))
    open(newunit=fid, file=trim(fin), status='old', form='formatted', action='read')
    do k = 1, nheader
       read(fid, *)
    end do
    do i = 1, nsig
       read(fid, *) t, re, im
       if (i == 1) t1 = t
       if (i == 2) dt = t - t1
       Correlator(i) = cmplx(re, im, fp)
    end do
    close(fid)

    call GetSpectralFunction(Correlator, dt, Spectrum, symm, hanned, fftformat)
    call write_signal('input.txt', Correlator)
    call write_signal('symm.txt', symm)
    call write_signal('hanned.txt', hanned)
    call write_signal('fftformat.txt', fftformat)

    nn = size(Spectrum, kind=int64)
    wmin = -pi / dt
    dw = 2 * pi / (dt * nn)
    open(newunit=fid, file=trim(fout), status='replace', form='formatted', action='write')
    do i = nn / 2 + 1, nn
       write(fid, '(2(SP,E13.6,1X))') wmin + (i - nn / 2 - 1) * dw, Spectrum(i)
    end do
    do i = 1, nn / 2
       write(fid, '(2(SP,E13.6,1X))') (i - 1) * dw, Spectrum(i)
    end do
    close(fid)
    write(output_unit, *) 'Simulation completed'
  end subroutine ComputeSpectrum

  !> index, real and imaginary part per row, list-directed (main.f90:1898-1950)
  subroutine write_signal(fname, x)
    character(len=*), intent(in) :: fname
    complex(fp), intent(in) :: x(:)
    integer :: fid
    integer(int64) :: i
    open(newunit=fid, file=fname, status='replace', form='formatted', action='write')
    do i = 1, size(x, kind=int64)
       write(fid, *) i, real(x(i)), aimag(x(i))
    end do
    close(fid)
  end subroutine write_signal

  !> Reads the heavy-quark parameters kspTol, M and the four Wilson
  !! coefficients (re, im) from arguments k0 .. k0 + 9
  subroutine read_quark_args(k0, kspTol, mass, wc)
    integer, intent(in) :: k0
    real(fp), intent(out) :: kspTol, mass
    complex(fp), intent(out) :: wc(4)
    integer :: k
    kspTol = rarg(k0)
    mass = rarg(k0 + 1)
    do k = 1, 4
       wc(k) = cmplx(rarg(k0 + 2 * k), rarg(k0 + 2 * k + 1), fp)
    end do
  end subroutine read_quark_args

  !> Header of the four-channel correlator files (main.f90:374-378)
  subroutine write_correlator_header(fid)
    integer, intent(in) :: fid
    write(fid, '(A1)', advance='no') 's'
    write(fid, '(1X,A11,1X,A11)', advance='no') 'Re(O1(1S0))', 'Im(O1(1S0))'
    write(fid, '(1X,A11,1X,A11)', advance='no') 'Re(O1(3S1))', 'Im(O1(3S1))'
    write(fid, '(1X,A11,1X,A11)', advance='no') 'Re(O8(1S0))', 'Im(O8(1S0))'
    write(fid, '(1X,A11,1X,A11)', advance='yes') 'Re(O8(3S1))', 'Im(O8(3S1))'
  end subroutine write_correlator_header

  !> One row of the norm file and of the correlator files (main.f90:385-417,
  !! 660-706); the _gconjg file fg only when present (units from newunit=
  !! are negative, so no unit value can mean "none")
  subroutine print_observables(s, HeavyField, fc, fn, fg)
    real(fp), intent(in) :: s
    type(NRQCDField), intent(in) :: HeavyField
    integer, intent(in) :: fc, fn
    integer, intent(in), optional :: fg
    complex(fp) :: c(4)
    write(fn, '(3(SP,E16.9,1X))') s, HeavyField%GetNorm_Quark(), HeavyField%GetNorm_AntiQ()
    c(1) = HeavyField%GetMesoncorrelator_O1_1S0_ZeroMomentum()
    c(2) = HeavyField%GetMesoncorrelator_O1_3S1_ZeroMomentum()
    c(3) = HeavyField%GetMesoncorrelator_O8_1S0_ZeroMomentum()
    c(4) = HeavyField%GetMesoncorrelator_O8_3S1_ZeroMomentum()
    write(fc, '(9(SP,E16.9,1X))') s, real(c(1)), aimag(c(1)), real(c(2)), aimag(c(2)), &
         real(c(3)), aimag(c(3)), real(c(4)), aimag(c(4))
    if (.not. present(fg)) return
    c(1) = HeavyField%GetMesoncorrelator_O1_1S0_ZeroMomentum_gconjg()
    c(2) = HeavyField%GetMesoncorrelator_O1_3S1_ZeroMomentum_gconjg()
    c(3) = HeavyField%GetMesoncorrelator_O8_1S0_ZeroMomentum_gconjg()
    c(4) = HeavyField%GetMesoncorrelator_O8_3S1_ZeroMomentum_gconjg()
    write(fg, '(9(SP,E16.9,1X))') s, real(c(1)), aimag(c(1)), real(c(2)), aimag(c(2)), &
         real(c(3)), aimag(c(3)), real(c(4)), aimag(c(4))
  end subroutine print_observables

  !> Mode 32: the point-source heavy field in a cold gauge field, all four
  !! correlators and both norms after every step (main.f90:282-522)
  subroutine MeasureHeavyQuarkoniumCorrelators_free()
    integer(int64) :: ext(3), seed, tsteps, it
    real(fp) :: spacings(0:3), smax, kspTol, mass
    complex(fp) :: wc(4)
    type(GaugeConfiguration) :: GaugeConf_t
    type(NRQCDField) :: HeavyField_t
    character(len=256) :: file_corr, file_norm
    integer :: k, fc, fn

    if (command_argument_count() < 22) call usage()
    do k = 1, 3
       ext(k) = iarg(1 + k)
    end do
    do k = 0, 3
       spacings(k) = rarg(5 + k)
    end do
    smax = rarg(9)
    seed = iarg(10)
    call read_quark_args(11, kspTol, mass, wc)
    call get_command_argument(21, file_corr)
    call get_command_argument(22, file_norm)

    call InitModule_Lattice(ext, spacings)
    call InitModule_Random(seed + ThisProc())
    call InitModule_NRQCD(mass, wc, 1._fp, kspTol)

    tsteps = nint(abs(smax) / spacings(0), int64)
    call GaugeConf_t%ColdInit()
    call HeavyField_t%InitSinglePoint(1_int64, 1_int64)

    open(newunit=fc, file=trim(file_corr), status='replace', form='formatted', action='write')
    open(newunit=fn, file=trim(file_norm), status='replace', form='formatted', action='write')
    call write_correlator_header(fc)
    write(fn, '(A1,1X,A5,1X,A5)') 's', 'Quark', 'Antiq'
    call print_observables(0._fp, HeavyField_t, fc, fn)
    do it = 1, tsteps
       call HeavyField_t%Update(GaugeConf_t, mass, wc)
       call GaugeConf_t%Update()
       call print_observables(spacings(0) * it, HeavyField_t, fc, fn)
       write(output_unit, '(F7.3,A1)') real(it) / tsteps * 100, '%'
    end do
    close(fc)
    close(fn)
    write(output_unit, '(A,I0)') 'Krylov iterations: ', TotalKrylovIterations
    call HeavyField_t%Destructor
    call GaugeConf_t%Deallocate
    call FinalizeModule_NRQCD()
    write(output_unit, *) 'Simulation completed'
  end subroutine MeasureHeavyQuarkoniumCorrelators_free

  !> Mode 33: as mode 32 from a thermal configuration (EquilibriumInit with
  !! Beta, nefieldinit, nequilibrium) instead of cold links (main.f90:18-280)
  subroutine MeasureHeavyQuarkoniumCorrelators_Equilibrium()
    integer(int64) :: ext(3), seed, tsteps, it, nefieldinit, nequilibrium
    real(fp) :: spacings(0:3), smax, beta, kspTol, mass
    complex(fp) :: wc(4)
    type(GaugeConfiguration) :: GaugeConf_t
    type(NRQCDField) :: HeavyField_t
    character(len=256) :: file_corr, file_norm, file_energy
    logical :: measure
    integer :: k, fc, fn

    if (command_argument_count() < 25) call usage()
    do k = 1, 3
       ext(k) = iarg(1 + k)
    end do
    do k = 0, 3
       spacings(k) = rarg(5 + k)
    end do
    smax = rarg(9)
    seed = iarg(10)
    beta = rarg(11)
    nefieldinit = iarg(12)
    nequilibrium = iarg(13)
    call read_quark_args(14, kspTol, mass, wc)
    call get_command_argument(24, file_corr)
    call get_command_argument(25, file_norm)
    ! a missing or zero flag: no energy measurement (the reference reads '(I1)')
    measure = .false.
    file_energy = ''
    if (command_argument_count() >= 26) then
       measure = iarg(26) /= 0
       if (measure) call get_command_argument(27, file_energy)
    end if

    call InitModule_Lattice(ext, spacings)
    call InitModule_Random(seed + ThisProc())
    call InitModule_NRQCD(mass, wc, 1._fp, kspTol)

    tsteps = nint(abs(smax) / spacings(0), int64)
    if (measure) then
       call GaugeConf_t%EquilibriumInit(beta, nefieldinit, nequilibrium, MeasureEnergy=measure, &
            filename=trim(file_energy))
    else
       call GaugeConf_t%EquilibriumInit(beta, nefieldinit, nequilibrium)
    end if
    call HeavyField_t%InitSinglePoint(1_int64, 1_int64)

    open(newunit=fc, file=trim(file_corr), status='replace', form='formatted', action='write')
    open(newunit=fn, file=trim(file_norm), status='replace', form='formatted', action='write')
    call write_correlator_header(fc)
    write(fn, '(A1,1X,A5,1X,A5)') 's', 'Quark', 'Antiq'
    call print_observables(0._fp, HeavyField_t, fc, fn)
    do it = 1, tsteps
       call HeavyField_t%Update(GaugeConf_t, mass, wc)
       call GaugeConf_t%Update()
       call print_observables(spacings(0) * it, HeavyField_t, fc, fn)
       write(output_unit, '(F7.3,A1)') real(it) / tsteps * 100, '%'
    end do
    close(fc)
    close(fn)
    write(output_unit, '(A,I0)') 'Krylov iterations: ', TotalKrylovIterations
    call HeavyField_t%Destructor
    call GaugeConf_t%Deallocate
    call FinalizeModule_NRQCD()
    write(output_unit, *) 'Simulation completed'
  end subroutine MeasureHeavyQuarkoniumCorrelators_Equilibrium

  !> Mode 31: correlators against the relative time s at one centre-of-mass
  !! time t (main.f90:524-830).  The gauge field is evolved to t; for
  !! it1 = 1 .. nint(smax/2/a0) a copy of it and the point source are evolved
  !! 2 it1 steps (CN, then the leapfrog), and the gauge field at t1 is stepped
  !! back once, so every row is centred on t.
  subroutine MeasureHeavyQuarkoniumCorrelators_oneT()
    integer(int64) :: ext(3), seed, t1steps, t2steps, it1, it2, iwork, nwork
    real(fp) :: spacings(0:3), t, smax, qs, amp, g, kspTol, mass
    complex(fp) :: wc(4)
    type(GaugeConfiguration) :: GaugeConf_t1, GaugeConf_t2
    type(NRQCDField) :: HeavyField_t1, HeavyField_t2
    character(len=256) :: file_corr, file_norm, file_gconjg
    integer :: k, fc, fn, fg

    if (command_argument_count() < 27) call usage()
    do k = 1, 3
       ext(k) = iarg(1 + k)
    end do
    do k = 0, 3
       spacings(k) = rarg(5 + k)
    end do
    t = rarg(9)
    smax = rarg(10)
    seed = iarg(11)
    qs = rarg(12)
    amp = rarg(13)
    g = rarg(14)
    call read_quark_args(15, kspTol, mass, wc)
    call get_command_argument(25, file_corr)
    call get_command_argument(26, file_norm)
    call get_command_argument(27, file_gconjg)

    call InitModule_Lattice(ext, spacings)
    call InitModule_Random(seed + ThisProc())
    call InitModule_NRQCD(mass, wc, 1._fp, kspTol)

    t1steps = nint(abs(smax / 2) / spacings(0), int64)
    nwork = t1steps * (t1steps + 1)
    call GaugeConf_t1%TransversePolarisedOccupiedInit_Box(qs, amp, g)
    call HeavyField_t1%InitSinglePoint(1_int64, 1_int64)
    if (nint(abs(t) / spacings(0)) > 0) &
         call GaugeConf_t1%Update(sign(1._fp, t), int(nint(abs(t) / spacings(0)), int64))

    open(newunit=fc, file=trim(file_corr), status='replace', form='formatted', action='write')
    open(newunit=fn, file=trim(file_norm), status='replace', form='formatted', action='write')
    open(newunit=fg, file=trim(file_gconjg), status='replace', form='formatted', action='write')
    call write_correlator_header(fc)
    write(fn, '(A1,1X,A5,1X,A5)') 's', 'Quark', 'Antiq'
    call write_correlator_header(fg)
    call print_observables(0._fp, HeavyField_t1, fc, fn, fg)

    iwork = 0
    do it1 = 1, t1steps
       t2steps = 2 * it1
       HeavyField_t2 = HeavyField_t1
       GaugeConf_t2 = GaugeConf_t1
       do it2 = 1, t2steps
          iwork = iwork + 1
          call HeavyField_t2%Update(GaugeConf_t2, mass, wc)
          call GaugeConf_t2%Update()
          write(output_unit, '(F7.3,A1)') real(iwork) / nwork * 100, '%'
       end do
       call print_observables(spacings(0) * t2steps, HeavyField_t2, fc, fn, fg)
       call GaugeConf_t1%Update(-1._fp)
    end do
    close(fc)
    close(fn)
    close(fg)
    write(output_unit, '(A,I0)') 'Krylov iterations: ', TotalKrylovIterations
    call HeavyField_t1%Destructor
    call HeavyField_t2%Destructor
    call GaugeConf_t1%Deallocate
    call GaugeConf_t2%Deallocate
    call FinalizeModule_NRQCD()
    write(output_unit, *) 'Simulation completed'
  end subroutine MeasureHeavyQuarkoniumCorrelators_oneT

  !> Mode 3: O1(3S1) correlator and propagator norms on the (t, s) grid; for
  !! every t1 the heavy field restarts from the point source and is evolved
  !! with the gauge field from t1 (CN first, then the leapfrog).
  subroutine MeasureHeavyQuarkoniumCorrelators()
    integer(int64) :: ext(3), seed, ismin, ismax, itmin, itmax, it1, it2, it2range, is, it
    integer(int64) :: nprogress, iprogress, ittest, istest
    real(fp) :: spacings(0:3), ds, tmin, tmax, smax, qs, amp, g, kspTol, mass, cre, cim
    real(fp) :: t1min, t1max, t1, t2, t, s, nan
    complex(fp) :: wc(4)
    complex(fp), allocatable :: correlator(:, :)
    real(fp), allocatable :: qnorm(:, :), anorm(:, :)
    type(GaugeConfiguration) :: GaugeConf_t1, GaugeConf_t2
    type(NRQCDField) :: HeavyField_t1, HeavyField_t2
    character(len=256) :: file_corr, file_norm
    logical :: on_grid
    integer :: k, fc, fn

    if (command_argument_count() < 26) call usage()
    do k = 1, 3
       ext(k) = iarg(1 + k)
    end do
    ds = rarg(5)
    spacings(0) = ds / 2
    do k = 1, 3
       spacings(k) = rarg(5 + k)
    end do
    tmin = rarg(9)
    tmax = rarg(10)
    smax = rarg(11)
    seed = iarg(12)
    qs = rarg(13)
    amp = rarg(14)
    g = rarg(15)
    kspTol = rarg(16)
    mass = rarg(17)
    do k = 1, 4
       cre = rarg(16 + 2 * k)
       cim = rarg(17 + 2 * k)
       wc(k) = cmplx(cre, cim, fp)
    end do
    call get_command_argument(26, file_corr)
    call get_command_argument(27, file_norm)
    if (len_trim(file_norm) == 0) file_norm = 'norm.txt'

    call InitModule_Lattice(ext, spacings)
    call InitModule_Random(seed + ThisProc())
    call InitModule_NRQCD(mass, wc, 1._fp, kspTol)

    nan = ieee_value(nan, ieee_quiet_nan)
    ismin = 0
    ismax = nint(smax / ds)
    itmin = nint(tmin / ds)
    itmax = nint(tmax / ds)
    allocate(correlator(ismin:ismax, itmin:itmax), qnorm(ismin:ismax, itmin:itmax), anorm(ismin:ismax, itmin:itmax))
    correlator = cmplx(nan, nan, fp)
    correlator(0, :) = cmplx(0, 1, fp)
    qnorm = nan
    qnorm(0, :) = 1
    anorm = qnorm

    call GaugeConf_t1%TransversePolarisedOccupiedInit_Box(qs, amp, g)
    call HeavyField_t1%InitSinglePoint(1_int64, 1_int64)

    t1min = tmin - smax / 2
    t1max = tmax
    ! evolve the gauge field from t = 0 to t1min
    if (nint(abs(t1min) / spacings(0)) > 0) &
         call GaugeConf_t1%Update(sign(1._fp, t1min), int(nint(abs(t1min) / spacings(0)), int64))

    it2range = nint(smax / spacings(0))
    nprogress = 0
    do it1 = nint(t1min / spacings(0), int64), nint(t1max / spacings(0), int64)
       do it2 = it1 + 1, it1 + it2range
          nprogress = nprogress + 1
          if ((it2 + it1) * spacings(0) / 2 > tmax) exit
       end do
    end do

    iprogress = 0
    do it1 = nint(t1min / spacings(0), int64), nint(t1max / spacings(0), int64)
       t1 = it1 * spacings(0)
       GaugeConf_t2 = GaugeConf_t1
       HeavyField_t2 = HeavyField_t1
       do it2 = it1 + 1, it1 + it2range
          iprogress = iprogress + 1
          write(output_unit, '(E16.9,A1)') real(iprogress) / max(nprogress, 1_int64) * 100, '%'
          ! the reference order: CN with U(t), E(t), then the leapfrog (main.f90:2329-2330)
          call HeavyField_t2%Update(GaugeConf_t2, mass, wc)
          call GaugeConf_t2%Update()
          t2 = it2 * spacings(0)
          t = (t2 + t1) / 2
          s = t2 - t1
          ittest = nint(t / ds)
          istest = nint(s / ds)
          on_grid = abs(ittest * ds - t) <= 1e-12_fp .and. abs(istest * ds - s) <= 1e-12_fp
          if (tmin <= t .and. t <= tmax .and. on_grid) then
             it = nint(t / ds)
             is = nint(s / ds)
             correlator(is, it) = HeavyField_t2%GetMesoncorrelator_O1_3S1_ZeroMomentum()
             qnorm(is, it) = HeavyField_t2%GetNorm_Quark()
             anorm(is, it) = HeavyField_t2%GetNorm_AntiQ()
          else if (t > tmax) then
             exit
          end if
       end do
       call GaugeConf_t1%Update()
    end do

    open(newunit=fc, file=trim(file_corr), status='replace', form='formatted', action='write')
    open(newunit=fn, file=trim(file_norm), status='replace', form='formatted', action='write')
    write(fc, '(A1,1X)', advance='no') 's'
    write(fn, '(A1,1X)', advance='no') 's'
    do it = itmin, itmax
       write(fc, '(A2,F0.6,1X,A2,F0.6,1X)', advance='no') 'Re', it * ds, 'Im', it * ds
       write(fn, '(A1,F0.6,1X,A1,F0.6,1X)', advance='no') 'Q', it * ds, 'A', it * ds
    end do
    write(fc, *)
    write(fn, *)
    do is = ismin, ismax
       write(fc, '(SP,E16.9,1X)', advance='no') is * ds
       write(fn, '(SP,E16.9,1X)', advance='no') is * ds
       do it = itmin, itmax
          write(fc, '(2(SP,E16.9,1X))', advance='no') real(correlator(is, it)), aimag(correlator(is, it))
          write(fn, '(2(SP,E16.9,1X))', advance='no') qnorm(is, it), anorm(is, it)
       end do
       write(fc, *)
       write(fn, *)
    end do
    close(fc)
    close(fn)
    write(output_unit, '(A,I0)') 'Krylov iterations: ', TotalKrylovIterations
    call HeavyField_t1%Destructor
    call HeavyField_t2%Destructor
    call GaugeConf_t1%Deallocate
    call GaugeConf_t2%Deallocate
    call FinalizeModule_NRQCD()
    write(output_unit, *) 'Simulation completed'
  end subroutine MeasureHeavyQuarkoniumCorrelators
end program nrqcd_run
