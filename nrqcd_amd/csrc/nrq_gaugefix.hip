// nrq_gaugefix.hip -- Coulomb gauge fixing and the transverse AA/EE
// correlators behind the gluon occupation f(p) = sqrt(AA EE)
// (gaugeconfiguration_su3.f90:1825-1972, 2269-2521; main.f90:3093-3189).
//
// Gauge fixing is Davies' Fourier-accelerated steepest descent.  Per
// iteration:
//   Delta(x) = sum_i skew(U_i(x - i) - U_i(x)),  skew(d) = (d - d^+)/2
// (not made traceless, as in the reference); stop when
// max_x ||Delta(x)||_F / 9 <= tol.  Otherwise Delta goes to momentum space,
// every mode with |p| > 1e-12 is multiplied by (p_max/|p|)^2, Delta comes back,
// is multiplied by alpha, and g(x) = exp(Delta(x)) transforms the links,
// U_i(x) <- g(x) U_i(x) g(x + i)^+.  E is not transformed (the reference's
// PerformGaugeTransformation touches only the links).  The nine components of
// Delta are one batched in-place hipFFT Z2Z pair; the reference's x2p/p2x
// scales (a1 a2 a3 forward, 1 / physical volume backward, xpfft.f90:134-136),
// the acceleration and alpha are one multiply in momentum space.
//
// The correlators take the 24 fields A^a_i(x) = GetGroupLog(U_i(x))^a / a_i
// (or E^a_i(x) / (a_i a0)) to momentum space in one batched transform and
// contract them per momentum with the transverse projector
// P_ij = Re[(delta_ij - conj(p_i) p_j / |p|^2) / 2] of the backward-derivative
// momentum (the reference's projector is a real function, lattice.f90:
// 1178-1195, so only the real part of the complex expression survives),
// divided by the physical volume and averaged over the generators.
//
// Group log: H = -i log U on the principal branch (eigenphases atan2(Im, Re)
// in (-pi, pi], su3.f90:318-334 via matrixoperations.f90:474-495), from the
// eigen-decomposition in nrq_grouplog.h.
#include "nrq_internal.h"
#include "nrq_grouplog.h"
#include <hipfft/hipfft.h>
#include <algorithm>
#include <cmath>
#include <vector>

namespace nrq {

const void *gauge_links(const nrq_gauge *h);
void *gauge_links_mut(nrq_gauge *h);

constexpr double GF_ZERO_TOL = 1e-12; // tolerances.f90
constexpr int GF_THREADS = 256;


template <typename R> __device__ __forceinline__ mat3<double> load_link_d(const cplx<R> *u, int64_t vol, int dir, int64_t s)
{
    mat3<double> m;
    const cplx<R> *p = u + (int64_t)dir * 9 * vol;
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        const cplx<R> a = *at32(p + (int64_t)e * vol, (uint32_t)s);
        m.m[e] = {(double)a.re, (double)a.im};
    }
    return m;
}

// Backward-derivative lattice momentum of momentum index s (lattice.f90:963-979)
// and its squared norm.
__device__ __forceinline__ double lattice_momentum(const geom &g, int64_t s, cplx<double> (&p)[3])
{
    const lsite q = make_site(g, s);
    const int n[3] = {q.x, q.y, q.z};
    const int ext[3] = {g.n0, g.n1, g.n2};
    double pn2 = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double arg = M_PI * (double)n[d] / (double)ext[d];
        double sn, cs;
        sincos(arg, &sn, &cs);
        const double mag = 2.0 * sn / g.a[d + 1];
        p[d] = {mag * cs, -mag * sn};
        pn2 += p[d].re * p[d].re + p[d].im * p[d].im;
    }
    return pn2;
}

// ---------------------------------------------------------------------------
// gauge fixing
// ---------------------------------------------------------------------------

// Delta(x) into nine complex planes; per-block max of ||Delta||_F / 9.
template <typename R>
__global__ __launch_bounds__(GF_THREADS) void k_gf_divergence(const cplx<R> *__restrict__ u, geom g,
                                                             cplx<double> *__restrict__ div, double *__restrict__ partial)
{
    __shared__ double sh[GF_THREADS / 64];
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double dev = 0;
    if (s < g.vol) {
        const lsite p = make_site(g, s);
        mat3<double> d;
        set_zero(d);
        for (int i = 0; i < 3; ++i) {
            const mat3<double> a = load_link_d(u, g.vol, i, site_dn(g, p, i).s);
            const mat3<double> b = load_link_d(u, g.vol, i, s);
            // skew(a - b) = ((a - b) - (a - b)^+) / 2, added entry by entry
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const cplx<double> x = a.m[3 * r + c] - b.m[3 * r + c];
                    const cplx<double> y = a.m[3 * c + r] - b.m[3 * c + r];
                    d.m[3 * r + c] = d.m[3 * r + c] + cplx<double>{0.5 * (x.re - y.re), 0.5 * (x.im + y.im)};
                }
        }
        double f = 0;
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            div[(int64_t)e * g.vol + s] = d.m[e];
            f += d.m[e].re * d.m[e].re + d.m[e].im * d.m[e].im;
        }
        dev = sqrt(f) / 9.0;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) dev = fmax(dev, __shfl_down(dev, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = dev;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0;
        for (int k = 0; k < GF_THREADS / 64; ++k) m = fmax(m, sh[k]);
        partial[blockIdx.x] = m;
    }
}

// max of n partials (one block)
__global__ __launch_bounds__(GF_THREADS) void k_gf_fold_max(const double *__restrict__ partial, int64_t n,
                                                           double *__restrict__ out)
{
    __shared__ double sh[GF_THREADS / 64];
    double m = 0;
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x) m = fmax(m, partial[k]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < GF_THREADS / 64; ++k) m = fmax(m, sh[k]);
        *out = m;
    }
}

// Momentum-space step: scale = fwd * bwd * alpha, times (p_max/|p|)^2 where |p| > tol.
__global__ void k_gf_accelerate(cplx<double> *__restrict__ div, geom g, double pmax, double scale)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.vol) return;
    cplx<double> p[3];
    const double pn = sqrt(lattice_momentum(g, s, p));
    double f = scale;
    if (pn > GF_ZERO_TOL) {
        const double r = pmax / pn;
        f *= r * r;
    }
#pragma unroll
    for (int e = 0; e < 9; ++e) div[(int64_t)e * g.vol + s] = f * div[(int64_t)e * g.vol + s];
}

// g = exp(Delta) at site s: exp(iQ) with Q = Hermitian part of -i Delta, its
// trace split off as a phase (the traceless part goes through the
// Cayley-Hamilton exponential).
__device__ __forceinline__ mat3<double> gf_group(const cplx<double> *__restrict__ div, int64_t vol, int64_t s)
{
    mat3<double> d;
#pragma unroll
    for (int e = 0; e < 9; ++e) d.m[e] = div[(int64_t)e * vol + s];
    mat3<double> q;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            // (-i d_rc + conj(-i d_cr)) / 2
            const cplx<double> a = d.m[3 * r + c], b = d.m[3 * c + r];
            q.m[3 * r + c] = {0.5 * (a.im + b.im), 0.5 * (-a.re + b.re)};
        }
    const double t = (q.m[0].re + q.m[4].re + q.m[8].re) / 3.0;
    q.m[0].re -= t;
    q.m[4].re -= t;
    q.m[8].re -= t;
    mat3<double> x = exp_i_hermitian(q);
    double sn, cs;
    sincos(t, &sn, &cs);
    const cplx<double> ph = {cs, sn};
#pragma unroll
    for (int e = 0; e < 9; ++e) x.m[e] = ph * x.m[e];
    return x;
}

// U_i(x) <- g(x) U_i(x) g(x + i)^+ for the three directions at x, with
// g = exp(Delta) formed here from Delta (gf_group) for x and its three
// forward neighbours, so no separate pass writes g: the exponentials are
// FP64 arithmetic inside a memory-bound kernel.
template <typename R>
__global__ __launch_bounds__(GF_THREADS) void k_gf_apply(cplx<R> *__restrict__ u, const cplx<double> *__restrict__ delta,
                                                         geom g)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.vol) return;
    const lsite p = make_site(g, s);
    const mat3<double> gx = gf_group(delta, g.vol, s);
    for (int i = 0; i < 3; ++i) {
        const int64_t sp = site_up(g, p, i).s;
        const mat3<double> gy = gf_group(delta, g.vol, sp);
        const mat3<double> v = mul_nd(mul(gx, load_link_d(u, g.vol, i, s)), gy);
        cplx<R> *b = u + (int64_t)i * 9 * g.vol;
#pragma unroll
        for (int e = 0; e < 9; ++e) *at32(b + (int64_t)e * g.vol, (uint32_t)s) = {(R)v.m[e].re, (R)v.m[e].im};
    }
}

// ---------------------------------------------------------------------------
// group log and correlators
// ---------------------------------------------------------------------------

// Cyclic Jacobi for a Hermitian 3x3 A: on return A is diagonal (real) and the
// columns of V are the eigenvectors.  Rotation in the (p, q) plane:
// J = diag(1, conj(e)) R, e = A_pq / |A_pq|, R = [[c, -s], [s, c]] with
// tan 2 theta = 2 |A_pq| / (A_pp - A_qq); A <- J^+ A J, V <- V J.
// 24 real fields (a + 8 i) as complex planes: AA (which = 0) or EE (which = 1)
template <typename R>
__global__ __launch_bounds__(GF_THREADS) void k_corr_fields(const cplx<R> *__restrict__ u, const R *__restrict__ ef, geom g,
                                                          int which, cplx<double> *__restrict__ fld)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.vol) return;
    for (int i = 0; i < 3; ++i) {
        double c[8];
        if (which == 0) {
            group_log_coords(load_link_d(u, g.vol, i, s), c);
            for (int a = 0; a < 8; ++a) c[a] = c[a] / g.a[i + 1];
        } else {
            for (int a = 0; a < 8; ++a) c[a] = (double)ef[((int64_t)i * 8 + a) * g.vol + s] / g.a[i + 1] / g.a[0];
        }
        for (int a = 0; a < 8; ++a) fld[(int64_t)(a + 8 * i) * g.vol + s] = {c[a], 0.0};
    }
}

// out[s] = Re sum_a sum_ij F^a_i P_ij conj(F^a_j) * scale, scale = fwd^2 / (8 V_phys),
// P_ij = (delta_ij - Re(conj(p_i) p_j) / |p|^2) / 2
__global__ void k_corr_project(const cplx<double> *__restrict__ fld, geom g, double scale, double *__restrict__ out)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.vol) return;
    cplx<double> p[3];
    const double pn2 = lattice_momentum(g, s, p);
    if (!(sqrt(pn2) >= GF_ZERO_TOL)) {
        out[s] = 0.0;
        return;
    }
    double rp[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) rp[i][j] = (p[i].re * p[j].re + p[i].im * p[j].im) / pn2; // Re(conj(p_i) p_j) / |p|^2
    double acc = 0;
    for (int a = 0; a < 8; ++a) {
        cplx<double> f[3];
        for (int i = 0; i < 3; ++i) f[i] = fld[(int64_t)(a + 8 * i) * g.vol + s];
        double t = 0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double gij = f[i].re * f[j].re + f[i].im * f[j].im; // Re(f_i conj(f_j))
                t += ((i == j ? 1.0 : 0.0) - rp[i][j]) * gij;
            }
        acc += 0.5 * t;
    }
    out[s] = acc * scale;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static int gf_fft_fail(hipfftResult r, const char *what)
{
    return fail(NRQ_ERR_HIP, "hipFFT %s failed (%d)", what, (int)r);
}

// Largest |p| of the backward-derivative momenta: |p|^2 is a sum of one term
// per direction, so its maximum is the sum of the per-direction maxima.
static double max_momentum(const geom &g)
{
    const int ext[3] = {g.n0, g.n1, g.n2};
    double sum = 0;
    for (int d = 0; d < 3; ++d) {
        double m = 0;
        for (int n = 0; n < ext[d]; ++n) {
            const double arg = M_PI * (double)n / (double)ext[d];
            const double mag = 2.0 * std::sin(arg) / g.a[d + 1];
            const double re = mag * std::cos(arg), im = -mag * std::sin(arg);
            m = std::max(m, re * re + im * im);
        }
        sum += m;
    }
    return std::sqrt(sum);
}

// Device buffers of one call, freed on every exit path.
struct gf_buffers {
    cplx<double> *fld = nullptr;
    double *partial = nullptr, *dmax = nullptr, *hmax = nullptr;
    ~gf_buffers()
    {
        if (fld) (void)hipFree(fld);
        if (partial) (void)hipFree(partial);
        if (dmax) (void)hipFree(dmax);
        if (hmax) (void)hipHostFree(hmax);
    }
};

// Batched in-place Z2Z plans, kept per (extents, batch) for the life of the
// process: plan creation costs milliseconds, a measurement repeated along a
// run should not pay it every time.
struct gf_plan_entry {
    int n0, n1, n2, batch;
    hipfftHandle plan;
};
static std::vector<gf_plan_entry> g_gf_plans;

static int gf_plan(const geom &g, int batch, hipfftHandle *out)
{
    for (const gf_plan_entry &e : g_gf_plans)
        if (e.n0 == g.n0 && e.n1 == g.n1 && e.n2 == g.n2 && e.batch == batch) {
            *out = e.plan;
            return hipfftSetStream(e.plan, stream()) == HIPFFT_SUCCESS ? NRQ_OK : fail(NRQ_ERR_HIP, "hipFFT set stream failed");
        }
    int dims[3] = {g.n2, g.n1, g.n0}; // row-major, x fastest
    hipfftHandle plan = 0;
    hipfftResult fr = hipfftPlanMany(&plan, 3, dims, nullptr, 1, (int)g.vol, nullptr, 1, (int)g.vol, HIPFFT_Z2Z, batch);
    if (fr != HIPFFT_SUCCESS) return gf_fft_fail(fr, "plan");
    fr = hipfftSetStream(plan, stream());
    if (fr != HIPFFT_SUCCESS) {
        hipfftDestroy(plan);
        return gf_fft_fail(fr, "set stream");
    }
    g_gf_plans.push_back({g.n0, g.n1, g.n2, batch, plan});
    *out = plan;
    return NRQ_OK;
}

template <typename R>
static int gaugefix_t(nrq_gauge *h, double tol, double alpha, int64_t maxit, int64_t *iterations, double *final_dev)
{
    const geom g = lat().g;
    const int64_t v = g.vol;
    const unsigned nblk = (unsigned)((v + GF_THREADS - 1) / GF_THREADS);
    gf_buffers b;
    if (hipMalloc((void **)&b.fld, sizeof(cplx<double>) * 9 * (size_t)v) != hipSuccess ||
        hipMalloc((void **)&b.partial, sizeof(double) * nblk) != hipSuccess ||
        hipMalloc((void **)&b.dmax, sizeof(double)) != hipSuccess ||
        hipHostMalloc((void **)&b.hmax, sizeof(double), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return fail(NRQ_ERR_ALLOC, "gauge fixing: device allocation failed (V=%lld)", (long long)v);
    }
    hipfftHandle plan = 0;
    if (int e = gf_plan(g, 9, &plan)) return e;
    cplx<R> *u = (cplx<R> *)gauge_links_mut(h);
    const double fwd = g.a[1] * g.a[2] * g.a[3];
    const double bwd = 1.0 / ((double)g.n0 * g.a[1] * (double)g.n1 * g.a[2] * (double)g.n2 * g.a[3]);
    const double pmax = max_momentum(g);
    int64_t done = 0;
    double dev = 0;
    for (;;) {
        hipLaunchKernelGGL(k_gf_divergence<R>, dim3(nblk), dim3(GF_THREADS), 0, stream(), u, g, b.fld, b.partial);
        hipLaunchKernelGGL(k_gf_fold_max, dim3(1), dim3(GF_THREADS), 0, stream(), b.partial, (int64_t)nblk, b.dmax);
        NRQ_HIP(hipGetLastError());
        NRQ_HIP(hipMemcpyAsync(b.hmax, b.dmax, sizeof(double), hipMemcpyDeviceToHost, stream()));
        NRQ_HIP(hipStreamSynchronize(stream()));
        dev = *b.hmax;
        if (!std::isfinite(dev)) return fail(NRQ_ERR_NOCONV, "gauge fixing: non-finite divergence");
        if (dev <= tol || done >= maxit) break;
        hipfftResult fr = hipfftExecZ2Z(plan, (hipfftDoubleComplex *)b.fld, (hipfftDoubleComplex *)b.fld, HIPFFT_FORWARD);
        if (fr != HIPFFT_SUCCESS) return gf_fft_fail(fr, "exec (forward)");
        hipLaunchKernelGGL(k_gf_accelerate, dim3(nblk), dim3(GF_THREADS), 0, stream(), b.fld, g, pmax, fwd * bwd * alpha);
        fr = hipfftExecZ2Z(plan, (hipfftDoubleComplex *)b.fld, (hipfftDoubleComplex *)b.fld, HIPFFT_BACKWARD);
        if (fr != HIPFFT_SUCCESS) return gf_fft_fail(fr, "exec (backward)");
        hipLaunchKernelGGL(k_gf_apply<R>, dim3(nblk), dim3(GF_THREADS), 0, stream(), u, b.fld, g);
        NRQ_HIP(hipGetLastError());
        ++done;
    }
    if (iterations) *iterations = done;
    if (final_dev) *final_dev = dev;
    return NRQ_OK;
}

int transverse_project(const cplx<double> *fld, double scale, double *out)
{
    const geom g = lat().g;
    const int64_t v = g.vol;
    double *dout = nullptr;
    if (hipMalloc((void **)&dout, sizeof(double) * (size_t)v) != hipSuccess) {
        (void)hipGetLastError();
        return fail(NRQ_ERR_ALLOC, "correlators: device allocation failed (V=%lld)", (long long)v);
    }
    const unsigned nblk = (unsigned)((v + GF_THREADS - 1) / GF_THREADS);
    hipLaunchKernelGGL(k_corr_project, dim3(nblk), dim3(GF_THREADS), 0, stream(), fld, g, scale, dout);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(double) * (size_t)v, hipMemcpyDeviceToHost, stream());
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    (void)hipFree(dout);
    return e == hipSuccess ? NRQ_OK : check_hip(e, "transverse projection");
}

template <typename R> static int correlators_t(const nrq_gauge *h, double *aa, double *ee)
{
    const geom g = lat().g;
    const int64_t v = g.vol;
    const unsigned nblk = (unsigned)((v + GF_THREADS - 1) / GF_THREADS);
    gf_buffers b;
    double *dout = nullptr;
    if (hipMalloc((void **)&b.fld, sizeof(cplx<double>) * 24 * (size_t)v) != hipSuccess ||
        hipMalloc((void **)&b.partial, sizeof(double) * (size_t)v) != hipSuccess) {
        (void)hipGetLastError();
        return fail(NRQ_ERR_ALLOC, "correlators: device allocation failed (V=%lld)", (long long)v);
    }
    dout = b.partial;
    hipfftHandle plan = 0;
    if (int e = gf_plan(g, 24, &plan)) return e;
    const double fwd = g.a[1] * g.a[2] * g.a[3];
    const double vphys = (double)g.n0 * g.a[1] * (double)g.n1 * g.a[2] * (double)g.n2 * g.a[3];
    const double scale = fwd * fwd / vphys / 8.0;
    const cplx<R> *u = (const cplx<R> *)gauge_links(h);
    const R *ef = (const R *)h->e;
    for (int which = 0; which < 2; ++which) {
        double *host = which == 0 ? aa : ee;
        if (!host) continue;
        hipLaunchKernelGGL(k_corr_fields<R>, dim3(nblk), dim3(GF_THREADS), 0, stream(), u, ef, g, which, b.fld);
        NRQ_HIP(hipGetLastError());
        hipfftResult fr = hipfftExecZ2Z(plan, (hipfftDoubleComplex *)b.fld, (hipfftDoubleComplex *)b.fld, HIPFFT_FORWARD);
        if (fr != HIPFFT_SUCCESS) return gf_fft_fail(fr, "exec (forward)");
        hipLaunchKernelGGL(k_corr_project, dim3(nblk), dim3(GF_THREADS), 0, stream(), b.fld, g, scale, dout);
        NRQ_HIP(hipGetLastError());
        NRQ_HIP(hipMemcpyAsync(host, dout, sizeof(double) * (size_t)v, hipMemcpyDeviceToHost, stream()));
        NRQ_HIP(hipStreamSynchronize(stream()));
    }
    return NRQ_OK;
}

} // namespace nrq

using namespace nrq;

extern "C" int nrq_gauge_coulomb_gaugefix(nrq_gauge *h, double tol, double alpha, int64_t max_iterations,
                                          int64_t *iterations, double *final_deviation)
{
    if (!lat().initialised) return fail(NRQ_ERR_STATE, "lattice: not initialised");
    if (!h) return fail(NRQ_ERR_ARG, "null gauge configuration");
    if (h->vol != lat().g.vol) return fail(NRQ_ERR_STATE, "gauge configuration does not match the lattice");
    if (!(tol >= 0) || !std::isfinite(alpha) || max_iterations < 0)
        return fail(NRQ_ERR_ARG, "gauge fixing: tolerance %g, alpha %g, max iterations %lld", tol, alpha,
                    (long long)max_iterations);
    if (int r = ensure_device()) return r;
    return h->precision == 64 ? gaugefix_t<double>(h, tol, alpha, max_iterations, iterations, final_deviation)
                              : gaugefix_t<float>(h, tol, alpha, max_iterations, iterations, final_deviation);
}

extern "C" int nrq_gauge_transverse_correlators(const nrq_gauge *h, double *aa, double *ee)
{
    if (!lat().initialised) return fail(NRQ_ERR_STATE, "lattice: not initialised");
    if (!h) return fail(NRQ_ERR_ARG, "null gauge configuration");
    if (h->vol != lat().g.vol) return fail(NRQ_ERR_STATE, "gauge configuration does not match the lattice");
    if (int r = ensure_device()) return r;
    return h->precision == 64 ? correlators_t<double>(h, aa, ee) : correlators_t<float>(h, aa, ee);
}

extern "C" int nrq_lattice_max_norm2_momentum(double *out)
{
    if (!lat().initialised) return fail(NRQ_ERR_STATE, "lattice: not initialised");
    if (!out) return fail(NRQ_ERR_ARG, "null argument");
    *out = max_momentum(lat().g);
    return NRQ_OK;
}
