// nrq_ic.hip -- transverse-polarised box initial condition on the GPU
// (gaugeconfiguration_su3.f90:906-951, 1263-1446 on one rank).
//
// The seeded noise comes from the host ranlxd stream in lattice order (the
// only sequential part); the 16 forward and 48 backward 3-D FFTs that the
// reference runs through MKL's cluster DFT (xpfft.f90) are two batched
// hipFFT Z2Z transforms on the whole lattice in HBM, with the reference's
// scales (forward a1 a2 a3, backward 1/Volume, xpfft.f90:134-136).
#include "nrq_internal.h"
#include <hipfft/hipfft.h>
#include <cmath>
#include <vector>

namespace nrq {

const void *gauge_links(const nrq_gauge *h);
void *gauge_links_mut(nrq_gauge *h);
void *gauge_efield(nrq_gauge *h);

__global__ void k_noise_to_complex(const double *__restrict__ r, cplx<double> *__restrict__ z, int64_t n)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) z[k] = {r[k], 0.0};
}

// Mode filling, gaugeconfiguration_su3.f90:1353-1381, with the backward
// derivative momentum (lattice.f90:963-979) and the Kaspar-Hebenstreit
// polarisation vectors (lattice.f90:1120-1176).  rk: 16 fields (a + 8 pol),
// ae: 48 fields (A: a + 8 i, E: 24 + a + 8 i).
__global__ void k_mode_fill(const cplx<double> *__restrict__ rk, cplx<double> *__restrict__ ae, geom g, double qs,
                            double amp, double coupling, double fwd_scale)
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
    // GetBoxOccupation, gaugeconfiguration_su3.f90:1455-1471
    const double f = pn < qs ? amp : coupling * coupling / 2.0;
    const double pa = sqrt(f / pn * 2.0);
    const double pe = sqrt(f * pn * 2.0); // times i
    for (int a = 0; a < 8; ++a) {
        cplx<double> r1 = rk[(int64_t)a * g.vol + s], r2 = rk[(int64_t)(a + 8) * g.vol + s];
        r1 = fwd_scale * r1;
        r2 = fwd_scale * r2;
        for (int i = 0; i < 3; ++i) {
            const cplx<double> mode = r1 * e1[i] + r2 * e2[i];
            ae[(int64_t)(a + 8 * i) * g.vol + s] = pa * mode;
            ae[(int64_t)(24 + a + 8 * i) * g.vol + s] = {-pe * mode.im, pe * mode.re};
        }
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

} // namespace nrq

using namespace nrq;

extern "C" int nrq_gauge_box_ic(nrq_gauge *h, double qs, double amp, double coupling)
{
    if (!lat().initialised) return fail(NRQ_ERR_STATE, "lattice: not initialised");
    if (!h || !h->u[0]) return fail(NRQ_ERR_ARG, "gauge: null or unallocated handle");
    if (h->vol != lat().g.vol) return fail(NRQ_ERR_STATE, "gauge: lattice changed since allocation");
    if (!std::isfinite(qs) || !std::isfinite(amp) || !std::isfinite(coupling))
        return fail(NRQ_ERR_ARG, "box IC: non-finite parameter");
    const geom g = lat().g;
    const int64_t v = g.vol;

    // random.f90:238-253 per site and polarisation: 8 complex normals from 16
    // uniforms, of which only the real parts are kept (:1334-1340).
    std::vector<double> noise((size_t)16 * (size_t)v);
    {
        double u[32];
        const double twopi = 2.0 * M_PI;
        for (int64_t s = 0; s < v; ++s) {
            rng_uniform(u, 32);
            for (int pol = 0; pol < 2; ++pol)
                for (int a = 0; a < 8; ++a) {
                    const double r1 = u[16 * pol + 2 * a], r2 = u[16 * pol + 2 * a + 1];
                    noise[(size_t)(a + 8 * pol) * (size_t)v + (size_t)s] = std::sqrt(-2.0 * std::log(r1)) * std::cos(twopi * r2);
                }
        }
    }
    double *dnoise = nullptr;
    cplx<double> *rk = nullptr, *ae = nullptr;
    hipfftHandle pf = 0, pb = 0;
    bool have_pf = false, have_pb = false;
    int status = NRQ_OK;
    auto cleanup = [&]() {
        if (have_pf) hipfftDestroy(pf);
        if (have_pb) hipfftDestroy(pb);
        if (dnoise) (void)hipFree(dnoise);
        if (rk) (void)hipFree(rk);
        if (ae) (void)hipFree(ae);
    };
    do {
        if (hipMalloc(&dnoise, sizeof(double) * 16 * (size_t)v) != hipSuccess ||
            hipMalloc(&rk, sizeof(cplx<double>) * 16 * (size_t)v) != hipSuccess ||
            hipMalloc(&ae, sizeof(cplx<double>) * 48 * (size_t)v) != hipSuccess) {
            status = fail(NRQ_ERR_ALLOC, "box IC: device allocation failed");
            break;
        }
        hipError_t he = hipMemcpyAsync(dnoise, noise.data(), sizeof(double) * 16 * (size_t)v, hipMemcpyHostToDevice, stream());
        if (he != hipSuccess) { status = check_hip(he, "box IC noise upload"); break; }
        const int64_t nn = 16 * v;
        hipLaunchKernelGGL(k_noise_to_complex, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, stream(), dnoise, rk, nn);
        int dims[3] = {g.n2, g.n1, g.n0}; // row-major: x fastest
        hipfftResult fr = hipfftPlanMany(&pf, 3, dims, nullptr, 1, (int)v, nullptr, 1, (int)v, HIPFFT_Z2Z, 16);
        if (fr != HIPFFT_SUCCESS) { status = fft_fail(fr, "plan (forward)"); break; }
        have_pf = true;
        hipfftSetStream(pf, stream());
        fr = hipfftExecZ2Z(pf, (hipfftDoubleComplex *)rk, (hipfftDoubleComplex *)rk, HIPFFT_FORWARD);
        if (fr != HIPFFT_SUCCESS) { status = fft_fail(fr, "exec (forward)"); break; }
        const double fwd = g.a[1] * g.a[2] * g.a[3];
        hipLaunchKernelGGL(k_mode_fill, dim3((unsigned)((v + 255) / 256)), dim3(256), 0, stream(), rk, ae, g, qs, amp, coupling, fwd);
        fr = hipfftPlanMany(&pb, 3, dims, nullptr, 1, (int)v, nullptr, 1, (int)v, HIPFFT_Z2Z, 48);
        if (fr != HIPFFT_SUCCESS) { status = fft_fail(fr, "plan (backward)"); break; }
        have_pb = true;
        hipfftSetStream(pb, stream());
        fr = hipfftExecZ2Z(pb, (hipfftDoubleComplex *)ae, (hipfftDoubleComplex *)ae, HIPFFT_BACKWARD);
        if (fr != HIPFFT_SUCCESS) { status = fft_fail(fr, "exec (backward)"); break; }
        const double vol_phys = (double)g.n0 * g.a[1] * (double)g.n1 * g.a[2] * (double)g.n2 * g.a[3];
        const double bwd = 1.0 / vol_phys;
        h->cur = 0;
        const dim3 grd((unsigned)((v + 255) / 256), 3);
        if (h->precision == 64)
            hipLaunchKernelGGL(k_ic_finalize<double>, grd, dim3(256), 0, stream(), ae, (cplx<double> *)h->u[0], (double *)h->e, g, bwd);
        else
            hipLaunchKernelGGL(k_ic_finalize<float>, grd, dim3(256), 0, stream(), ae, (cplx<float> *)h->u[0], (float *)h->e, g, bwd);
        he = hipGetLastError();
        if (he != hipSuccess) { status = check_hip(he, "box IC kernels"); break; }
        he = hipStreamSynchronize(stream());
        if (he != hipSuccess) { status = check_hip(he, "box IC synchronise"); break; }
    } while (0);
    cleanup();
    h->steps_since_reunit = 0;
    return status;
}
