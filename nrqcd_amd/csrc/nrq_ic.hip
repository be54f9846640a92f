// nrq_ic.hip -- transverse-polarised initial conditions on the GPU
// (gaugeconfiguration_su3.f90:906-951 box occupation, 1263-1446
// TransversePolarisedOccupiedInit, 959-1255 its _specificSender form, on one
// rank), for the box occupation or any occupation per momentum.
//
// The seeded noise comes from the host ranlxd stream in lattice order (the
// only sequential part).  Per-site form: 16 real noise fields go to momentum
// space in one batched hipFFT (the reference's 16 forward x2p).  Specific-
// sender form: the complex amplitudes are drawn directly in momentum space
// and made Hermitian afterwards.  Then one kernel fills the 48 components of
// A and E per momentum, one batched backward hipFFT (the reference's 48 p2x)
// and one kernel exponentiate the links.  The reference's scales: forward
// a1 a2 a3, backward 1/Volume (xpfft.f90:134-136).
#include "nrq_internal.h"
#include <hipfft/hipfft.h>
#include <cmath>
#include <vector>

namespace nrq {

const void *gauge_links(const nrq_gauge *h);
void *gauge_links_mut(nrq_gauge *h);
void *gauge_efield(nrq_gauge *h);

// Occupation per momentum: an array by lattice index (device), or the box
// GetBoxOccupation(|p|, qs, amp, coupling), gaugeconfiguration_su3.f90:1455-1471.
struct ic_occupation {
    const double *f = nullptr;
    double qs = 0, amp = 0, coupling = 0;
};

// Mode filling with the backward-derivative momentum (lattice.f90:963-979) and
// the Kaspar-Hebenstreit polarisation vectors (lattice.f90:1120-1176).
// rk: 16 fields (a + 8 pol); ae: 48 fields (A: a + 8 i, E: 24 + a + 8 i).
// Per-site form (gaugeconfiguration_su3.f90:1353-1381): rk is the unscaled
// forward transform of real noise, A = sqrt(2 f/|p|) (r1 e1 + r2 e2),
// E = i sqrt(2 f |p|) (...).  Specific-sender form (:1036-1076): rk holds the
// complex amplitudes, A = sqrt(V f/|p|/2) (...), E = i sqrt(V f |p|/2) (...),
// V the physical volume.
__global__ void k_mode_fill(const cplx<double> *__restrict__ rk, cplx<double> *__restrict__ ae, geom g,
                            ic_occupation occ, int sender, double fwd_scale, double vphys)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.vol) return;
    const int64_t n[3] = {s % g.n0, (s / g.n0) % g.n1, s / ((int64_t)g.n0 * g.n1)};
    const int32_t ext[3] = {g.n0, g.n1, g.n2};
    cplx<double> p[3];
    double pabs2[3], pn2 = 0;
    for (int d = 0; d < 3; ++d) {
        const double arg = M_PI * (double)n[d] / (double)ext[d];
        double sn, cs;
        sincos(arg, &sn, &cs);
        const double mag = 2.0 * sn / g.a[d + 1];
        p[d] = {mag * cs, -mag * sn};
        pabs2[d] = p[d].re * p[d].re + p[d].im * p[d].im;
        pn2 += pabs2[d];
    }
    const double pn = sqrt(pn2);
    const double tol = 1e-12; // tolerances.f90:23,26
    if (!(pn > tol)) {
        for (int f = 0; f < 48; ++f) ae[(int64_t)f * g.vol + s] = {0.0, 0.0};
        return;
    }
    cplx<double> e1[3], e2[3];
    if (sqrt(pabs2[0]) > tol) {
        const double n12 = sqrt(pabs2[0] + pabs2[1]);
        e1[0] = (-1.0 / n12) * p[1];
        e1[1] = (1.0 / n12) * p[0];
        e1[2] = {0.0, 0.0};
        const double den = 1.0 / (pn * n12);
        e2[0] = den * cmul(p[0], p[2]);
        e2[1] = den * cmul(p[1], p[2]);
        e2[2] = {-n12 / pn, 0.0};
    } else {
        const double n23 = sqrt(pabs2[1] + pabs2[2]);
        e1[0] = {0.0, 0.0};
        e1[1] = (1.0 / n23) * p[2];
        e1[2] = (-1.0 / n23) * p[1];
        e2[0] = {1.0, 0.0};
        e2[1] = {0.0, 0.0};
        e2[2] = {0.0, 0.0};
    }
    const double f = occ.f ? occ.f[s] : (pn < occ.qs ? occ.amp : occ.coupling * occ.coupling / 2.0);
    const double pa = sender ? sqrt(vphys * f / pn / 2.0) : sqrt(f / pn * 2.0);
    const double pe = sender ? sqrt(vphys * f * pn / 2.0) : sqrt(f * pn * 2.0); // times i
    const double rs = sender ? 1.0 : fwd_scale;
    for (int a = 0; a < 8; ++a) {
        cplx<double> r1 = rk[(int64_t)a * g.vol + s], r2 = rk[(int64_t)(a + 8) * g.vol + s];
        r1 = rs * r1;
        r2 = rs * r2;
        for (int i = 0; i < 3; ++i) {
            const cplx<double> mode = r1 * e1[i] + r2 * e2[i];
            ae[(int64_t)(a + 8 * i) * g.vol + s] = pa * mode;
            ae[(int64_t)(24 + a + 8 * i) * g.vol + s] = {-pe * mode.im, pe * mode.re};
        }
    }
}

// SymmetriseInPspace (gaugeconfiguration_su3.f90:1163-1255) of the 48
// components: walking the lattice indices in order, the entry at the smaller
// of p, -p is kept and the other becomes its conjugate; a self-conjugate
// momentum keeps its real part.  Thread s writes only entries with
// -s < s (from the kept entry at -s, which no thread writes) or s = -s.
__global__ void k_ic_symmetrise(cplx<double> *__restrict__ ae, geom g)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.vol) return;
    const int64_t x = s % g.n0, y = (s / g.n0) % g.n1, z = s / ((int64_t)g.n0 * g.n1);
    const int64_t t = (x == 0 ? 0 : g.n0 - x) + (int64_t)g.n0 * ((y == 0 ? 0 : g.n1 - y) + (int64_t)g.n1 * (z == 0 ? 0 : g.n2 - z));
    if (t < s) {
        for (int f = 0; f < 48; ++f) {
            const cplx<double> v = ae[(int64_t)f * g.vol + t];
            ae[(int64_t)f * g.vol + s] = {v.re, -v.im};
        }
    } else if (t == s) {
        for (int f = 0; f < 48; ++f) ae[(int64_t)f * g.vol + s].im = 0.0;
    }
}

// gaugeconfiguration_su3.f90:1431-1443: U_i = exp(i a_i Re A^a T^a),
// E^a_i = a0 a_i Re E^a.
template <typename R>
__global__ void k_ic_finalize(const cplx<double> *__restrict__ ae, cplx<R> *__restrict__ u, R *__restrict__ ef, geom g,
                              double bwd_scale)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (s >= g.vol) return;
    double al[8];
    for (int a = 0; a < 8; ++a) al[a] = (ae[(int64_t)(a + 8 * i) * g.vol + s].re * bwd_scale) * g.a[i + 1];
    mat3<double> x = exp_i_hermitian(algebra_matrix(al));
    for (int e = 0; e < 9; ++e) u[((int64_t)i * 9 + e) * g.vol + s] = {(R)x.m[e].re, (R)x.m[e].im};
    for (int a = 0; a < 8; ++a)
        ef[((int64_t)i * 8 + a) * g.vol + s] = (R)(((ae[(int64_t)(24 + a + 8 * i) * g.vol + s].re * bwd_scale) * g.a[0]) * g.a[i + 1]);
}

static int fft_fail(hipfftResult r, const char *what)
{
    return fail(NRQ_ERR_HIP, "hipFFT error %d in %s", (int)r, what);
}

// Initial condition with occupation occ_host[V] (by lattice index) or, when it
// is null, the box (qs, amp, coupling); the momentum-space correlators of A
// and E go to aa, ee when they are not null (gaugeconfiguration_su3.f90:
// 1390-1421 and 1105-1136).
static int transverse_ic(nrq_gauge *h, const double *occ_host, double qs, double amp, double coupling, bool sender,
                         double *aa, double *ee)
{
    if (!lat().initialised) return fail(NRQ_ERR_STATE, "lattice: not initialised");
    if (!h || !h->u[0]) return fail(NRQ_ERR_ARG, "gauge: null or unallocated handle");
    if (h->vol != lat().g.vol) return fail(NRQ_ERR_STATE, "gauge: lattice changed since allocation");
    if (!occ_host && (!std::isfinite(qs) || !std::isfinite(amp) || !std::isfinite(coupling)))
        return fail(NRQ_ERR_ARG, "box IC: non-finite parameter");
    if (!rng_ready()) return fail(NRQ_ERR_STATE, "initial condition: random module not initialised");
    const geom g = lat().g;
    const int64_t v = g.vol;
    if (occ_host)
        for (int64_t s = 0; s < v; ++s)
            if (!(occ_host[s] >= 0.0) || !std::isfinite(occ_host[s]))
                return fail(NRQ_ERR_ARG, "initial condition: occupation %g at lattice index %lld", occ_host[s],
                            (long long)(s + 1));

    // Noise in lattice order.  Per-site form (random.f90:238-253,
    // gaugeconfiguration_su3.f90:1334-1340): per site and polarisation 8
    // complex normals from 16 uniforms, real parts kept.  Specific-sender form
    // (:1036-1054): per site with |p| > 1e-12 and generator 8 complex normals
    // from 16 uniforms, the first two kept.
    const double twopi = 2.0 * M_PI;
    std::vector<cplx<double>> noise((size_t)16 * (size_t)v, cplx<double>{0.0, 0.0});
    if (!sender) {
        double u[32];
        for (int64_t s = 0; s < v; ++s) {
            rng_uniform(u, 32);
            for (int pol = 0; pol < 2; ++pol)
                for (int a = 0; a < 8; ++a) {
                    const double r1 = u[16 * pol + 2 * a], r2 = u[16 * pol + 2 * a + 1];
                    noise[(size_t)(a + 8 * pol) * (size_t)v + (size_t)s] = {std::sqrt(-2.0 * std::log(r1)) * std::cos(twopi * r2), 0.0};
                }
        }
    } else {
        double u[16];
        for (int64_t s = 0; s < v; ++s) {
            const int64_t n[3] = {s % g.n0, (s / g.n0) % g.n1, s / ((int64_t)g.n0 * g.n1)};
            const int ext[3] = {g.n0, g.n1, g.n2};
            double pn2 = 0;
            for (int d = 0; d < 3; ++d) {
                const double mag = 2.0 * std::sin(M_PI * (double)n[d] / (double)ext[d]) / g.a[d + 1];
                pn2 += mag * mag;
            }
            if (!(std::sqrt(pn2) > 1e-12)) continue;
            for (int a = 0; a < 8; ++a) {
                rng_uniform(u, 16);
                for (int pol = 0; pol < 2; ++pol) {
                    const double m = std::sqrt(-2.0 * std::log(u[2 * pol]));
                    noise[(size_t)(a + 8 * pol) * (size_t)v + (size_t)s] = {m * std::cos(twopi * u[2 * pol + 1]),
                                                                            m * std::sin(twopi * u[2 * pol + 1])};
                }
            }
        }
    }
    double *docc = nullptr;
    cplx<double> *rk = nullptr, *ae = nullptr;
    hipfftHandle pf = 0, pb = 0;
    bool have_pf = false, have_pb = false;
    int status = NRQ_OK;
    auto cleanup = [&]() {
        if (have_pf) hipfftDestroy(pf);
        if (have_pb) hipfftDestroy(pb);
        if (docc) (void)hipFree(docc);
        if (rk) (void)hipFree(rk);
        if (ae) (void)hipFree(ae);
    };
    do {
        if (hipMalloc(&rk, sizeof(cplx<double>) * 16 * (size_t)v) != hipSuccess ||
            hipMalloc(&ae, sizeof(cplx<double>) * 48 * (size_t)v) != hipSuccess ||
            (occ_host && hipMalloc(&docc, sizeof(double) * (size_t)v) != hipSuccess)) {
            (void)hipGetLastError();
            status = fail(NRQ_ERR_ALLOC, "initial condition: device allocation failed");
            break;
        }
        hipError_t he = hipMemcpyAsync(rk, noise.data(), sizeof(cplx<double>) * 16 * (size_t)v, hipMemcpyHostToDevice, stream());
        if (he == hipSuccess && occ_host)
            he = hipMemcpyAsync(docc, occ_host, sizeof(double) * (size_t)v, hipMemcpyHostToDevice, stream());
        if (he != hipSuccess) { status = check_hip(he, "initial condition upload"); break; }
        int dims[3] = {g.n2, g.n1, g.n0}; // row-major: x fastest
        hipfftResult fr;
        if (!sender) {
            fr = hipfftPlanMany(&pf, 3, dims, nullptr, 1, (int)v, nullptr, 1, (int)v, HIPFFT_Z2Z, 16);
            if (fr != HIPFFT_SUCCESS) { status = fft_fail(fr, "plan (forward)"); break; }
            have_pf = true;
            hipfftSetStream(pf, stream());
            fr = hipfftExecZ2Z(pf, (hipfftDoubleComplex *)rk, (hipfftDoubleComplex *)rk, HIPFFT_FORWARD);
            if (fr != HIPFFT_SUCCESS) { status = fft_fail(fr, "exec (forward)"); break; }
        }
        const double fwd = g.a[1] * g.a[2] * g.a[3];
        const double vol_phys = (double)g.n0 * g.a[1] * (double)g.n1 * g.a[2] * (double)g.n2 * g.a[3];
        ic_occupation occ;
        occ.f = docc;
        occ.qs = qs;
        occ.amp = amp;
        occ.coupling = coupling;
        const unsigned nb = (unsigned)((v + 255) / 256);
        hipLaunchKernelGGL(k_mode_fill, dim3(nb), dim3(256), 0, stream(), rk, ae, g, occ, sender ? 1 : 0, fwd, vol_phys);
        if (sender) hipLaunchKernelGGL(k_ic_symmetrise, dim3(nb), dim3(256), 0, stream(), ae, g);
        he = hipGetLastError();
        if (he != hipSuccess) { status = check_hip(he, "initial condition mode filling"); break; }
        // gaugeconfiguration_su3.f90:1390-1421: sum_a Re[X_i P_ij conj(X_j)] / V / 8
        if (aa && (status = transverse_project(ae, 1.0 / vol_phys / 8.0, aa)) != NRQ_OK) break;
        if (ee && (status = transverse_project(ae + (size_t)24 * (size_t)v, 1.0 / vol_phys / 8.0, ee)) != NRQ_OK) break;
        fr = hipfftPlanMany(&pb, 3, dims, nullptr, 1, (int)v, nullptr, 1, (int)v, HIPFFT_Z2Z, 48);
        if (fr != HIPFFT_SUCCESS) { status = fft_fail(fr, "plan (backward)"); break; }
        have_pb = true;
        hipfftSetStream(pb, stream());
        fr = hipfftExecZ2Z(pb, (hipfftDoubleComplex *)ae, (hipfftDoubleComplex *)ae, HIPFFT_BACKWARD);
        if (fr != HIPFFT_SUCCESS) { status = fft_fail(fr, "exec (backward)"); break; }
        const double bwd = 1.0 / vol_phys;
        h->cur = 0;
        const dim3 grd(nb, 3);
        if (h->precision == 64)
            hipLaunchKernelGGL(k_ic_finalize<double>, grd, dim3(256), 0, stream(), ae, (cplx<double> *)h->u[0], (double *)h->e, g, bwd);
        else
            hipLaunchKernelGGL(k_ic_finalize<float>, grd, dim3(256), 0, stream(), ae, (cplx<float> *)h->u[0], (float *)h->e, g, bwd);
        he = hipGetLastError();
        if (he != hipSuccess) { status = check_hip(he, "initial condition kernels"); break; }
        he = hipStreamSynchronize(stream());
        if (he != hipSuccess) { status = check_hip(he, "initial condition synchronise"); break; }
    } while (0);
    cleanup();
    h->steps_since_reunit = 0;
    return status;
}

} // namespace nrq

using namespace nrq;

extern "C" int nrq_gauge_box_ic(nrq_gauge *h, double qs, double amp, double coupling)
{
    return transverse_ic(h, nullptr, qs, amp, coupling, false, nullptr, nullptr);
}

extern "C" int nrq_gauge_box_ic_correlators(nrq_gauge *h, double qs, double amp, double coupling, double *aa,
                                            double *ee)
{
    return transverse_ic(h, nullptr, qs, amp, coupling, false, aa, ee);
}

extern "C" int nrq_gauge_transverse_init(nrq_gauge *h, const double *occupation, int specific_sender, double *aa,
                                         double *ee)
{
    if (!occupation) return fail(NRQ_ERR_ARG, "initial condition: null occupation");
    return transverse_ic(h, occupation, 0.0, 0.0, 0.0, specific_sender != 0, aa, ee);
}
