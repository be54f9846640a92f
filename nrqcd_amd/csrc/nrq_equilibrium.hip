// nrq_equilibrium.hip -- EquilibriumInit (gaugeconfiguration_su3.f90:380-814)
// on the GPU: thermal electric fields projected onto the Gauss law, then
// leapfrog equilibration, repeated nefieldinit times from cold links.
//
// Per outer iteration:
//   E^a_k(x) = Re(sigma r^a) a0, sigma = sqrt(4/beta), r: 8 complex normals
//     per site and direction from 16 uniforms of the host ranlxd stream in
//     lattice order (:455-466);
//   G(x) = [sum_i (E_i(x) - U_i(x-i)^+ E_i(x-i) U_i(x-i)) / a_i
//           - sum_a T^a Tr(T^a diag(rho(x)) / (a1 a2 a3))] / a0   (GetG, :781-813),
//   deviation = max_{x,a} 2 |Tr(T^a G(x))|                  (GetGdeviation, :724-779);
//   while deviation > tol, from the previous configuration (Jacobi sweep):
//     E_k(x) <- coordinates of kappa a0 (U_k(x) G(x) U_k(x)^+ G(x+k) - G(x)) + E_k(x),
//     kappa = 0.01 (:480-507; the product of G(x) and G(x+k) is the
//     reference's, taken literally);
//   nequilibrium leapfrog steps of step width 1 (:696-698).
// The default tolerance is the reference's single-precision literal 1E-3.
// k_eq_gauss writes G and per-block maxima of the deviation, k_eq_fold_max
// folds them; k_eq_project writes the new E into the other of two buffers.
// The sweeps run in batches without host synchronisation: every kernel of a
// sweep first reads a device flag that the previous sweep's fold set once
// the deviation reached the tolerance, and returns at once if it is set, so
// exactly the reference's number of sweeps changes E.
#include "nrq_internal.h"
#include <algorithm>
#include <cmath>
#include <vector>

namespace nrq {

const void *gauge_links(const nrq_gauge *h);

constexpr int EQ_THREADS = 256;
constexpr double EQ_KAPPA = 0.01;
constexpr int64_t EQ_MAX_PROJECTIONS = 1000000; // guard: the reference loops without bound

template <typename R> __device__ __forceinline__ mat3<double> eq_link(const cplx<R> *u, int64_t vol, int dir, int64_t s)
{
    mat3<double> m;
    const cplx<R> *p = u + (int64_t)dir * 9 * vol;
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        const cplx<R> a = p[(int64_t)e * vol + s];
        m.m[e] = {(double)a.re, (double)a.im};
    }
    return m;
}

template <typename R> __device__ __forceinline__ mat3<double> eq_efield(const R *ef, int64_t vol, int dir, int64_t s)
{
    double al[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) al[a] = (double)ef[((int64_t)dir * 8 + a) * vol + s];
    return algebra_matrix(al);
}

// Loop state on the device: done once the deviation is <= tol, the number of
// sweeps that changed E, and the latest deviation.
struct eq_state {
    int done;
    int pad;
    long long sweeps;
    double dev;
};

// G(x) into gm[e V + s] (9 complex) and the block maxima of 2 |Tr(T^a G)|;
// nothing when st->done (st may be null: always)
template <typename R>
__global__ __launch_bounds__(EQ_THREADS) void k_eq_gauss(const cplx<R> *__restrict__ u, const R *__restrict__ ef,
                                                        const double *__restrict__ rho, geom g,
                                                        cplx<double> *__restrict__ gm, double *__restrict__ bmax,
                                                        const eq_state *__restrict__ st)
{
    if (st && st->done) return;
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double dev = 0.0;
    if (s < g.vol) {
        const int64_t V = g.vol;
        const lsite p = make_site(g, s);
        mat3<double> G;
        set_zero(G);
        for (int i = 0; i < 3; ++i) {
            const int64_t m = site_dn(g, p, i).s;
            const mat3<double> Um = eq_link(u, V, i, m);
            const mat3<double> t = mul(mul_dn(Um, eq_efield(ef, V, i, m)), Um);
            const mat3<double> Ex = eq_efield(ef, V, i, s);
            const double ia = 1.0 / g.a[i + 1];
#pragma unroll
            for (int e = 0; e < 9; ++e)
                G.m[e] = {G.m[e].re + (Ex.m[e].re - t.m[e].re) * ia, G.m[e].im + (Ex.m[e].im - t.m[e].im) * ia};
        }
        if (rho) {
            // sum_a T^a Tr(T^a C), C = diag(rho) / (a1 a2 a3): Tr(T^a C) is
            // non-zero for the two diagonal generators only
            const double vcell = g.a[1] * g.a[2] * g.a[3];
            double c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            const double r0 = rho[3 * s] / vcell, r1 = rho[3 * s + 1] / vcell, r2 = rho[3 * s + 2] / vcell;
            c[2] = 0.5 * (r0 - r1);
            c[7] = 0.28867513459481288225457439025098 * (r0 + r1 - 2.0 * r2);
            const mat3<double> q = algebra_matrix(c);
#pragma unroll
            for (int e = 0; e < 9; ++e) G.m[e] = {G.m[e].re - q.m[e].re, G.m[e].im - q.m[e].im};
        }
        const double ia0 = 1.0 / g.a[0];
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            G.m[e] = {G.m[e].re * ia0, G.m[e].im * ia0};
            gm[(int64_t)e * V + s] = G.m[e];
        }
        double tr[8];
        two_re_trace_gen(G, tr);
#pragma unroll
        for (int a = 0; a < 8; ++a) dev = fmax(dev, fabs(tr[a]));
    }
    __shared__ double sh[EQ_THREADS / 64];
    for (int off = 32; off > 0; off >>= 1) dev = fmax(dev, __shfl_down(dev, off, 64));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = dev;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int k = 0; k < EQ_THREADS / 64; ++k) m = fmax(m, sh[k]);
        bmax[blockIdx.x] = m;
    }
}

// Deviation = max of the block maxima.  count = 0: the deviation of the
// starting configuration (sets done, sweeps = 0); count = 1: the deviation
// after a sweep, unless done (sweeps + 1, then done = dev <= tol).
__global__ __launch_bounds__(EQ_THREADS) void k_eq_fold_max(const double *__restrict__ bmax, int nb, eq_state *__restrict__ st,
                                                           double tol, int count)
{
    if (count && st->done) return;
    double m = 0.0;
    for (int b = threadIdx.x; b < nb; b += EQ_THREADS) m = fmax(m, bmax[b]);
    __shared__ double sh[EQ_THREADS / 64];
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < EQ_THREADS / 64; ++k) t = fmax(t, sh[k]);
        st->dev = t;
        st->sweeps = count ? st->sweeps + 1 : 0;
        st->done = t <= tol ? 1 : 0;
    }
}

// E_new,k(x) = coordinates of kdt (U_k(x) G(x) U_k(x)^+ G(x+k) - G(x)) + E_k(x)
template <typename R>
__global__ __launch_bounds__(EQ_THREADS) void k_eq_project(const cplx<R> *__restrict__ u, const R *__restrict__ ef,
                                                          const cplx<double> *__restrict__ gm, geom g, double kdt,
                                                          R *__restrict__ ef_new, const eq_state *__restrict__ st)
{
    if (st->done) return;
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.vol) return;
    const int64_t V = g.vol;
    const lsite p = make_site(g, s);
    mat3<double> G;
#pragma unroll
    for (int e = 0; e < 9; ++e) G.m[e] = gm[(int64_t)e * V + s];
    for (int k = 0; k < 3; ++k) {
        const int64_t n = site_up(g, p, k).s;
        mat3<double> Gn;
#pragma unroll
        for (int e = 0; e < 9; ++e) Gn.m[e] = gm[(int64_t)e * V + n];
        const mat3<double> U = eq_link(u, V, k, s);
        const mat3<double> A = mul(mul_nd(mul(U, G), U), Gn);
        const mat3<double> E = eq_efield(ef, V, k, s);
        mat3<double> term;
#pragma unroll
        for (int e = 0; e < 9; ++e)
            term.m[e] = {kdt * (A.m[e].re - G.m[e].re) + E.m[e].re, kdt * (A.m[e].im - G.m[e].im) + E.m[e].im};
        double c[8];
        two_re_trace_gen(term, c);
#pragma unroll
        for (int a = 0; a < 8; ++a) ef_new[((int64_t)k * 8 + a) * V + s] = (R)c[a];
    }
}

// Sweeps launched per batch between host checks of the loop state.
constexpr int EQ_BATCH = 64;

template <typename R>
static int equilibrium_t(nrq_gauge *h, double beta, int64_t nefieldinit, int64_t nequilibrium, double tol,
                         const double *rho_host, double *energies, double *deviations, int64_t *projections)
{
    const geom g = lat().g;
    const int64_t V = g.vol;
    const unsigned nb = (unsigned)((V + EQ_THREADS - 1) / EQ_THREADS);
    struct bufs {
        cplx<double> *gm = nullptr;
        double *bmax = nullptr, *rho = nullptr;
        eq_state *st = nullptr, *hst = nullptr;
        void *ealt = nullptr;
        ~bufs()
        {
            if (gm) (void)hipFree(gm);
            if (bmax) (void)hipFree(bmax);
            if (st) (void)hipFree(st);
            if (rho) (void)hipFree(rho);
            if (ealt) (void)hipFree(ealt);
            if (hst) (void)hipHostFree(hst);
        }
    } b;
    const size_t ebytes = sizeof(R) * 24 * (size_t)V;
    if (hipMalloc((void **)&b.gm, sizeof(cplx<double>) * 9 * (size_t)V) != hipSuccess ||
        hipMalloc((void **)&b.bmax, sizeof(double) * nb) != hipSuccess ||
        hipMalloc((void **)&b.st, sizeof(eq_state)) != hipSuccess || hipMalloc(&b.ealt, ebytes) != hipSuccess ||
        (rho_host && hipMalloc((void **)&b.rho, sizeof(double) * 3 * (size_t)V) != hipSuccess) ||
        hipHostMalloc((void **)&b.hst, sizeof(eq_state), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return fail(NRQ_ERR_ALLOC, "EquilibriumInit: device allocation failed (V=%lld)", (long long)V);
    }
    if (rho_host) NRQ_HIP(hipMemcpyAsync(b.rho, rho_host, sizeof(double) * 3 * (size_t)V, hipMemcpyHostToDevice, stream()));
    if (int r = nrq_gauge_cold(h)) return r;
    R *ebuf[2] = {(R *)h->e, (R *)b.ealt};
    // deviation of the configuration in h->e (starts the loop state); the
    // links are looked up at every use, since the leapfrog swaps their buffers
    auto start = [&]() -> int {
        hipLaunchKernelGGL(k_eq_gauss<R>, dim3(nb), dim3(EQ_THREADS), 0, stream(), (const cplx<R> *)gauge_links(h),
                           (const R *)h->e, b.rho, g, b.gm, b.bmax, (const eq_state *)nullptr);
        hipLaunchKernelGGL(k_eq_fold_max, dim3(1), dim3(EQ_THREADS), 0, stream(), b.bmax, (int)nb, b.st, tol, 0);
        NRQ_HIP(hipGetLastError());
        return NRQ_OK;
    };
    auto read_state = [&]() -> int {
        NRQ_HIP(hipMemcpyAsync(b.hst, b.st, sizeof(eq_state), hipMemcpyDeviceToHost, stream()));
        NRQ_HIP(hipStreamSynchronize(stream()));
        return NRQ_OK;
    };
    const double sigma = std::sqrt(4.0 / beta);
    const double a0 = g.a[0];
    const double twopi = 2.0 * M_PI;
    std::vector<R> ehost((size_t)24 * (size_t)V);
    for (int64_t it = 0; it < nefieldinit; ++it) {
        // thermal E: per site (lattice order) and direction, 8 complex
        // normals from 16 uniforms, real parts (random.f90:238-253)
        double uu[16];
        for (int64_t s = 0; s < V; ++s)
            for (int k = 0; k < 3; ++k) {
                rng_uniform(uu, 16);
                for (int a = 0; a < 8; ++a) {
                    const double re = std::sqrt(-2.0 * std::log(uu[2 * a])) * std::cos(twopi * uu[2 * a + 1]);
                    ehost[((size_t)k * 8 + a) * (size_t)V + (size_t)s] = (R)((sigma * re) * a0);
                }
            }
        NRQ_HIP(hipMemcpyAsync(h->e, ehost.data(), ebytes, hipMemcpyHostToDevice, stream()));
        if (int r = start()) return r;
        if (int r = read_state()) return r;
        // sweep j (from 0) reads buffer j % 2 and writes the other; sweeps
        // after the one that reached the tolerance do nothing
        const cplx<R> *u = (const cplx<R> *)gauge_links(h); // fixed during the projection
        int64_t launched = 0;
        while (!b.hst->done) {
            if (launched >= EQ_MAX_PROJECTIONS)
                return fail(NRQ_ERR_NOCONV, "EquilibriumInit: Gauss-law projection above %g after %lld sweeps (deviation %g)",
                            tol, (long long)EQ_MAX_PROJECTIONS, b.hst->dev);
            for (int j = 0; j < EQ_BATCH; ++j, ++launched) {
                const R *src = ebuf[launched & 1];
                R *dst = ebuf[(launched + 1) & 1];
                hipLaunchKernelGGL(k_eq_project<R>, dim3(nb), dim3(EQ_THREADS), 0, stream(), u, src, b.gm, g, EQ_KAPPA * a0,
                                   dst, (const eq_state *)b.st);
                hipLaunchKernelGGL(k_eq_gauss<R>, dim3(nb), dim3(EQ_THREADS), 0, stream(), u, (const R *)dst, b.rho, g,
                                   b.gm, b.bmax, (const eq_state *)b.st);
                hipLaunchKernelGGL(k_eq_fold_max, dim3(1), dim3(EQ_THREADS), 0, stream(), b.bmax, (int)nb, b.st, tol, 1);
            }
            NRQ_HIP(hipGetLastError());
            if (int r = read_state()) return r;
        }
        const int64_t nproj = b.hst->sweeps;
        if (nproj & 1) // the final E is in the second buffer
            NRQ_HIP(hipMemcpyAsync(h->e, b.ealt, ebytes, hipMemcpyDeviceToDevice, stream()));
        if (nequilibrium > 0)
            if (int r = nrq_gauge_leapfrog(h, 1.0, nequilibrium)) return r;
        if (energies)
            if (int r = nrq_gauge_energy(h, &energies[it])) return r;
        if (deviations) {
            if (int r = start()) return r;
            if (int r = read_state()) return r;
            deviations[it] = b.hst->dev;
        }
        if (projections) projections[it] = nproj;
    }
    return NRQ_OK;
}

} // namespace nrq

using namespace nrq;

extern "C" int nrq_gauge_equilibrium_init(nrq_gauge *h, double beta, int64_t nefieldinit, int64_t nequilibrium,
                                          double tolerance, const double *charge_density, double *energies,
                                          double *deviations, int64_t *projections)
{
    if (!lat().initialised) return fail(NRQ_ERR_STATE, "lattice: not initialised");
    if (!h || !h->u[0]) return fail(NRQ_ERR_ARG, "gauge: null or unallocated handle");
    if (h->vol != lat().g.vol) return fail(NRQ_ERR_STATE, "gauge: lattice changed since allocation");
    if (!(beta > 0) || !std::isfinite(beta)) return fail(NRQ_ERR_ARG, "EquilibriumInit: beta must be positive");
    if (nefieldinit < 0 || nequilibrium < 0) return fail(NRQ_ERR_ARG, "EquilibriumInit: negative iteration count");
    if (!rng_ready()) return fail(NRQ_ERR_STATE, "EquilibriumInit: random module not initialised");
    if (!(tolerance > 0)) tolerance = (double)1e-3f; // the reference's default, a single-precision literal
    if (int r = ensure_device()) return r;
    return h->precision == 64
               ? equilibrium_t<double>(h, beta, nefieldinit, nequilibrium, tolerance, charge_density, energies,
                                       deviations, projections)
               : equilibrium_t<float>(h, beta, nefieldinit, nequilibrium, tolerance, charge_density, energies,
                                      deviations, projections);
}
