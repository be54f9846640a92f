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
!!   32 the same as 3 from a cold (free) gauge field (main.f90:282-522)
!! Reals are read list-directed (the reference reads '(F10.13)', which
!! accepts the same inputs); an optional trailing argument 'fp32' in mode 2
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
     call MeasureHeavyQuarkoniumCorrelators(.false.)
  case (32)
     call MeasureHeavyQuarkoniumCorrelators(.true.)
  case default
     call usage()
  end select

contains

  subroutine usage()
    write(output_unit, '(A)') 'usage: nrqcd_run 1|2|3|32 <positional parameters of the reference mode>'
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

  !> Modes 3 / 32: O1(3S1) correlator and propagator norms on the (t, s)
  !! grid; for every t1 the heavy field restarts from the point source and is
  !! evolved with the gauge field from t1 (CN first, then the leapfrog).
  subroutine MeasureHeavyQuarkoniumCorrelators(cold)
    logical, intent(in) :: cold
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

    if (cold) then
       call GaugeConf_t1%ColdInit()
    else
       call GaugeConf_t1%TransversePolarisedOccupiedInit_Box(qs, amp, g)
    end if
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
