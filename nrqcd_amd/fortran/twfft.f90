!> twfft module (twfft.f90:14-106): 1-D transforms between time and
!! frequency on the GPU, and the spectral function of ComputeSpectrum
!! (main.f90:1865-2100) for whole correlators.
module twfft
  use, intrinsic :: iso_c_binding
  use, intrinsic :: iso_fortran_env, only: int64
  use precision, only: fp
  use nrq_capi
  implicit none
  private
  public :: t2w, w2t, GetSpectralFunction

contains

  !> t2w, twfft.f90:73-105: data <- dt sum_j data_j exp(+2 pi i j l / n)
  subroutine t2w(data_, dt)
    complex(fp), intent(inout), target, contiguous :: data_(:)
    real(fp), intent(in) :: dt
    if (size(data_) == 0) return
    call nrq_check(nrq_t2w(c_loc(data_), int(size(data_), c_int64_t), dt), 't2w')
  end subroutine t2w

  !> w2t, twfft.f90:33-66: data <- sum_l data_l exp(-2 pi i j l / n) / (dt n)
  subroutine w2t(data_, dt)
    complex(fp), intent(inout), target, contiguous :: data_(:)
    real(fp), intent(in) :: dt
    if (size(data_) == 0) return
    call nrq_check(nrq_w2t(c_loc(data_), int(size(data_), c_int64_t), dt), 'w2t')
  end subroutine w2t

  !> The spectrum 2 Im t2w(Symm2FFTformat(Hann(Symmetrisation(C)))) of the
  !! correlator C(0), C(dt), ... (main.f90:1918-1961), size 2 size(C) - 2, in
  !! transform order.  The optional arrays receive the intermediate signals
  !! (symmetrised, windowed, reordered).
  subroutine GetSpectralFunction(Correlator, dt, Spectrum, Symmetrised, Windowed, Reordered)
    complex(fp), intent(in), target, contiguous :: Correlator(:)
    real(fp), intent(in) :: dt
    real(fp), allocatable, target, intent(out) :: Spectrum(:)
    complex(fp), allocatable, intent(out), optional :: Symmetrised(:), Windowed(:), Reordered(:)
    complex(fp), allocatable, target :: stages(:, :)
    integer(int64) :: n, nn
    type(c_ptr) :: sp
    n = size(Correlator, kind=int64)
    nn = max(2 * n - 2, 0_int64)
    allocate(Spectrum(nn))
    sp = c_null_ptr
    if (present(Symmetrised) .or. present(Windowed) .or. present(Reordered)) then
       allocate(stages(nn, 3))
       sp = c_loc(stages)
    end if
    call nrq_check(nrq_spectral_functions(c_loc(Correlator), n, 1_c_int64_t, dt, c_loc(Spectrum), sp), &
         'GetSpectralFunction')
    if (present(Symmetrised)) Symmetrised = stages(:, 1)
    if (present(Windowed)) Windowed = stages(:, 2)
    if (present(Reordered)) Reordered = stages(:, 3)
  end subroutine GetSpectralFunction
end module twfft
