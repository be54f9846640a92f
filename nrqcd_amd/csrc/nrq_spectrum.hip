// nrq_spectrum.hip -- spectral functions of heavy-quarkonium correlators
// (ComputeSpectrum, main.f90:1865-2100, with ApplyHannWindow,
// windowing.f90:11-25) and the 1-D time/frequency transforms t2w and w2t
// (twfft.f90:33-105).
//
// A correlator C(t_m), t_m = m dt, m = 0 .. n-1, becomes a signal of
// N = 2n - 2 points (0-based index k):
//   symmetrised  s_k = -conj(C(t_{n-1-k}))  for k = 0 .. n-2,
//                s_k = C(t_{k-n+1})         for k = n-1 .. 2n-3
// (main.f90:1970-1992: negative times carry -conj of the positive ones;
// C(t_{n-1}) appears at t = -(n-1) dt only, where the window vanishes),
//   windowed     h_k = s_k sin^2(pi k / N)  (windowing.f90:11-25),
//   rotated      f_j = h_{(j + N/2) mod N}  (t = 0 first, main.f90:2000-2016),
// transformed backward with scale dt, F_l = dt sum_j f_j exp(+2 pi i j l / N)
// (MKL's DFTI backward transform, twfft.f90:73-105), and the spectrum is
// rho_l = 2 Im F_l in transform order (l = 0: omega = 0; the writer orders the
// frequencies, main.f90:2066-2099).
//
// One kernel builds the rotated windowed signals of a whole batch of
// correlators, one batched in-place hipFFT Z2Z transforms them, one kernel
// takes 2 dt Im.  All three are streaming passes (HBM-bound); the plans are
// kept per (N, batch).
#include "nrq_internal.h"
#include <hipfft/hipfft.h>
#include <cmath>
#include <vector>

namespace nrq {

constexpr int SPEC_THREADS = 256;

// Writes symmetrised entry k of member b: window, rotation, stages.
__device__ __forceinline__ void spec_put(cplx<double> s, int64_t k, int64_t b, int64_t N, cplx<double> *__restrict__ f,
                                         cplx<double> *__restrict__ stages)
{
    const double pi = 3.14159265358979323846;
    // sin(pi (i-1) / N)^2 with the reference's evaluation order
    const double sn = sin((pi * (double)k) / (double)N);
    const double w = sn * sn;
    const cplx<double> h = {s.re * w, s.im * w};
    int64_t j = k + N / 2;
    if (j >= N) j -= N;
    f[b * N + j] = h;
    if (stages) {
        stages[(b * 3 + 0) * N + k] = s;
        stages[(b * 3 + 1) * N + k] = h;
        stages[(b * 3 + 2) * N + j] = h;
    }
}

// Builds f (rotated windowed signals, batch x N) from the correlators
// c (batch x n), one thread per correlator value, which lands at two places:
// C(t_m) at k = m + n - 1 (m <= n - 2) and -conj(C(t_m)) at k = n - 1 - m
// (m >= 1).  Optionally the three intermediate signals of every member,
// stages[(b * 3 + q) * N + k] for q = symmetrised, windowed, rotated.
__global__ __launch_bounds__(SPEC_THREADS) void k_spec_build(const cplx<double> *__restrict__ c, int64_t n,
                                                            int64_t total, cplx<double> *__restrict__ f,
                                                            cplx<double> *__restrict__ stages)
{
    const int64_t N = 2 * n - 2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = i / n, m = i - b * n;
        const cplx<double> v = c[i];
        if (m <= n - 2) spec_put(v, m + n - 1, b, N, f, stages);
        if (m >= 1) spec_put({-v.re, v.im}, n - 1 - m, b, N, f, stages);
    }
}

// rho = 2 Im(dt F)
__global__ __launch_bounds__(SPEC_THREADS) void k_spec_finish(const cplx<double> *__restrict__ F, int64_t total,
                                                             double dt, double *__restrict__ rho)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        rho[i] = 2.0 * (F[i].im * dt);
}

__global__ __launch_bounds__(SPEC_THREADS) void k_spec_scale(cplx<double> *__restrict__ F, int64_t total, double s)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const cplx<double> v = F[i];
        F[i] = {v.re * s, v.im * s};
    }
}

static int spec_fft_fail(hipfftResult r, const char *what)
{
    return fail(NRQ_ERR_HIP, "hipFFT %s failed (%d)", what, (int)r);
}

// 1-D batched in-place Z2Z plans, kept per (length, batch) for the life of
// the process.
struct spec_plan_entry {
    int64_t len, batch;
    hipfftHandle plan;
};
static std::vector<spec_plan_entry> g_spec_plans;

static int spec_plan(int64_t len, int64_t batch, hipfftHandle *out)
{
    for (const spec_plan_entry &e : g_spec_plans)
        if (e.len == len && e.batch == batch) {
            *out = e.plan;
            const hipfftResult r = hipfftSetStream(e.plan, stream());
            return r == HIPFFT_SUCCESS ? NRQ_OK : spec_fft_fail(r, "set stream");
        }
    int dims[1] = {(int)len};
    hipfftHandle plan = 0;
    hipfftResult r = hipfftPlanMany(&plan, 1, dims, nullptr, 1, (int)len, nullptr, 1, (int)len, HIPFFT_Z2Z, (int)batch);
    if (r != HIPFFT_SUCCESS) return spec_fft_fail(r, "plan");
    r = hipfftSetStream(plan, stream());
    if (r != HIPFFT_SUCCESS) {
        hipfftDestroy(plan);
        return spec_fft_fail(r, "set stream");
    }
    g_spec_plans.push_back({len, batch, plan});
    *out = plan;
    return NRQ_OK;
}

static scratch_buf g_spec_signal, g_spec_work, g_spec_rho, g_spec_stages;

static unsigned spec_blocks(int64_t total)
{
    const int64_t b = (total + SPEC_THREADS - 1) / SPEC_THREADS;
    return (unsigned)std::min<int64_t>(std::max<int64_t>(b, 1), 8192);
}

// Largest transform handed to hipFFT (lengths and batch counts are int
// there): at most 2^30 complex points per call, 16 GiB of signal.
constexpr int64_t SPEC_MAX_ELEMENTS = (int64_t)1 << 30;

static int spectral_functions(const double *signals, int64_t n, int64_t batch, double dt, double *spectra,
                              double *stages)
{
    if (!signals || !spectra) return fail(NRQ_ERR_ARG, "spectral function: null signal or spectrum");
    if (n < 2) return fail(NRQ_ERR_ARG, "spectral function: the correlator needs at least 2 time points (got %lld)", (long long)n);
    if (batch < 1) return fail(NRQ_ERR_ARG, "spectral function: batch must be positive (got %lld)", (long long)batch);
    if (!std::isfinite(dt) || dt == 0.0) return fail(NRQ_ERR_ARG, "spectral function: time step must be finite and non-zero");
    const int64_t N = 2 * n - 2;
    if (N > SPEC_MAX_ELEMENTS / batch)
        return fail(NRQ_ERR_ARG, "spectral function: %lld signals of %lld points exceed %lld points per call",
                    (long long)batch, (long long)N, (long long)SPEC_MAX_ELEMENTS);
    if (int r = ensure_device()) return r;
    const int64_t total = N * batch;
    if (int r = g_spec_signal.reserve(sizeof(cplx<double>) * (size_t)(n * batch))) return r;
    if (int r = g_spec_work.reserve(sizeof(cplx<double>) * (size_t)total)) return r;
    if (int r = g_spec_rho.reserve(sizeof(double) * (size_t)total)) return r;
    if (stages)
        if (int r = g_spec_stages.reserve(sizeof(cplx<double>) * 3 * (size_t)total)) return r;
    hipfftHandle plan;
    if (int r = spec_plan(N, batch, &plan)) return r;
    cplx<double> *c = (cplx<double> *)g_spec_signal.ptr, *f = (cplx<double> *)g_spec_work.ptr;
    cplx<double> *st = stages ? (cplx<double> *)g_spec_stages.ptr : nullptr;
    NRQ_HIP(hipMemcpyAsync(c, signals, sizeof(cplx<double>) * (size_t)(n * batch), hipMemcpyHostToDevice, stream()));
    hipLaunchKernelGGL(k_spec_build, dim3(spec_blocks(n * batch)), dim3(SPEC_THREADS), 0, stream(), c, n, n * batch, f,
                       st);
    NRQ_HIP(hipGetLastError());
    const hipfftResult r = hipfftExecZ2Z(plan, (hipfftDoubleComplex *)f, (hipfftDoubleComplex *)f, HIPFFT_BACKWARD);
    if (r != HIPFFT_SUCCESS) return spec_fft_fail(r, "backward transform");
    double *rho = (double *)g_spec_rho.ptr;
    hipLaunchKernelGGL(k_spec_finish, dim3(spec_blocks(total)), dim3(SPEC_THREADS), 0, stream(), f, total, dt, rho);
    NRQ_HIP(hipGetLastError());
    NRQ_HIP(hipMemcpyAsync(spectra, rho, sizeof(double) * (size_t)total, hipMemcpyDeviceToHost, stream()));
    if (stages)
        NRQ_HIP(hipMemcpyAsync(stages, st, sizeof(cplx<double>) * 3 * (size_t)total, hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    return NRQ_OK;
}

// In-place 1-D transform of n complex points with the given direction and
// scale (twfft.f90:33-105).
static int time_frequency(double *data, int64_t n, int direction, double scale)
{
    if (!data) return fail(NRQ_ERR_ARG, "t2w/w2t: null data");
    if (n < 1 || n > SPEC_MAX_ELEMENTS) return fail(NRQ_ERR_ARG, "t2w/w2t: bad length %lld", (long long)n);
    if (int r = ensure_device()) return r;
    if (int r = g_spec_work.reserve(sizeof(cplx<double>) * (size_t)n)) return r;
    hipfftHandle plan = 0;
    if (n > 1)
        if (int r = spec_plan(n, 1, &plan)) return r;
    cplx<double> *f = (cplx<double> *)g_spec_work.ptr;
    NRQ_HIP(hipMemcpyAsync(f, data, sizeof(cplx<double>) * (size_t)n, hipMemcpyHostToDevice, stream()));
    if (n > 1) { // a one-point transform is the identity
        const hipfftResult r = hipfftExecZ2Z(plan, (hipfftDoubleComplex *)f, (hipfftDoubleComplex *)f, direction);
        if (r != HIPFFT_SUCCESS) return spec_fft_fail(r, "transform");
    }
    hipLaunchKernelGGL(k_spec_scale, dim3(spec_blocks(n)), dim3(SPEC_THREADS), 0, stream(), f, n, scale);
    NRQ_HIP(hipGetLastError());
    NRQ_HIP(hipMemcpyAsync(data, f, sizeof(cplx<double>) * (size_t)n, hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    return NRQ_OK;
}

} // namespace nrq

using namespace nrq;

extern "C" int nrq_spectral_functions(const double *signals, int64_t n, int64_t batch, double dt, double *spectra,
                                      double *stages)
{
    return spectral_functions(signals, n, batch, dt, spectra, stages);
}

extern "C" int nrq_t2w(double *data, int64_t n, double dt)
{
    if (!std::isfinite(dt)) return fail(NRQ_ERR_ARG, "t2w: time step must be finite");
    return time_frequency(data, n, HIPFFT_BACKWARD, dt);
}

extern "C" int nrq_w2t(double *data, int64_t n, double dt)
{
    if (!std::isfinite(dt) || dt == 0.0) return fail(NRQ_ERR_ARG, "w2t: time step must be finite and non-zero");
    return time_frequency(data, n, HIPFFT_FORWARD, 1.0 / (dt * (double)n));
}
