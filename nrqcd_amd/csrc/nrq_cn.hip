// nrq_cn.hip -- NRQCD heavy-quark Crank-Nicolson step (nrqcd.f90:1053-1168)
// with a matrix-free covariant-Laplacian stencil and a batched restarted
// GMRES(m) with classical Gram-Schmidt that replaces PETSc KSP.
//
// Storage.  For the kinetic Hamiltonian H = -c1 D^2/2M the 6x6 propagator
// P(d, j, x), d = 2(colour-1) + spin-1 (the C2CS = U (x) 1_2 ordering of
// nrqcd.f90:452-467), splits into independent colour vectors: one per
// (propagator, column j, spin s).  The handle stores 24 such fields
// (quark: 0..11, antiquark: 12..23; field = 12 p + 2 j + s), each as
// psi[field][colour][site] complex double.
//
// Operator.  With K psi(x) = sum_i [psi(x) - (U_i(x) psi(x+i) +
// U_i(x-i)^+ psi(x-i))/2] / a_i^2 and alpha = -i (StepWidth a0/2) c1/M,
// the reference builds S = 1 + alpha K (nrqcd.f90:1203-1242), forms
// rhs = S psi (:1112) and solves S^+ psi' = rhs (:1097,:1123) with
// S^+ = 1 + conj(alpha) K.  The antiquark uses M -> -M, c -> conj(c)
// (:1130-1161).  Every field is one right-hand side of a batched solve that
// shares each link load across the fields of a chunk.
#include "nrq_krylov.h"
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

namespace nrq {

const void *gauge_links(const nrq_gauge *h);

constexpr int NFIELDS = NFIELDS_PER_HEAVY;

struct cn_config {
    bool initialised = false;
    double mass = 1, stepwidth = 1;
    std::complex<double> c[4];
    krylov_params solver;
};
static cn_config g_cn;

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Block-level sum across the block's waves; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    v = wsum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += sh[k];
    return t;
}

__global__ void k_fold8(const double *__restrict__ partial, int nblk, double *__restrict__ out)
{
    const int q = threadIdx.x;
    if (q >= 8) return;
    double s = 0;
    for (int b = 0; b < nblk; ++b) s += partial[(int64_t)q * nblk + b];
    out[q] = s;
}

__global__ void k_set_point(cplx<double> *psi, int64_t vol, int first_field, int64_t s, double v)
{
    // InitSinglePoint: P(d, j) = delta_dj v; field (j, spin) holds colour c with
    // d = 2c + spin, so the non-zero is colour c = (j - spin)/2 when j - spin is even.
    const int t = threadIdx.x;
    if (t >= 12) return;
    const int j = t / 2, sp = t % 2;
    if ((j - sp) % 2 != 0) return;
    const int c = (j - sp) / 2;
    psi[((int64_t)(first_field + t) * 3 + c) * vol + s] = {v, 0.0};
}

// Meson correlators per site (nrqcd.f90:1484-1661) from the 24 colour fields.
// Vertex M = X (x) Y with X in {1, T^a} (reference generator table) and
// Y in {1, sigma_k} (su2.f90:68-88, no 1/2):
//   Tr(A^+ M Q M) = sum conj(A_pq) M_pr Q_rt M_tq.
__device__ __forceinline__ cplx<double> gen_elem(int a, int r, int c)
{
    // T^a_{rc}, a = 0..7, reference convention (see nrq_su3.h)
    const double h = 0.5, s = 0.28867513459481288225457439025098;
    switch (a) {
    case 0: return {(r == 0 && c == 1) || (r == 1 && c == 0) ? h : 0.0, 0.0};
    case 1: return {0.0, (r == 0 && c == 1) ? h : ((r == 1 && c == 0) ? -h : 0.0)};
    case 2: return {(r == c) ? (r == 0 ? h : (r == 1 ? -h : 0.0)) : 0.0, 0.0};
    case 3: return {(r == 0 && c == 2) || (r == 2 && c == 0) ? h : 0.0, 0.0};
    case 4: return {0.0, (r == 0 && c == 2) ? h : ((r == 2 && c == 0) ? -h : 0.0)};
    case 5: return {(r == 1 && c == 2) || (r == 2 && c == 1) ? h : 0.0, 0.0};
    case 6: return {0.0, (r == 1 && c == 2) ? h : ((r == 2 && c == 1) ? -h : 0.0)};
    default: return {(r == c) ? (r == 2 ? -2.0 * s : s) : 0.0, 0.0};
    }
}
__device__ __forceinline__ cplx<double> pauli_elem(int k, int r, int c)
{
    // sigma_k (k = 1..3) in the reference's convention, k = 0: identity
    switch (k) {
    case 0: return {r == c ? 1.0 : 0.0, 0.0};
    case 1: return {r != c ? 1.0 : 0.0, 0.0};
    case 2: return {0.0, (r == 0 && c == 1) ? 1.0 : ((r == 1 && c == 0) ? -1.0 : 0.0)};
    default: return {r == c ? (r == 0 ? 1.0 : -1.0) : 0.0, 0.0};
    }
}

__global__ __launch_bounds__(REDUCE_THREADS) void k_corr(const cplx<double> *__restrict__ psi, int64_t vol,
                                                         double *__restrict__ partial)
{
    // acc: O1_1S0, O1_3S1, O8_1S0, O8_3S1 (re, im)
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < vol; s += (int64_t)gridDim.x * blockDim.x) {
        cplx<double> Q[36], A[36];
        for (int j = 0; j < 6; ++j)
            for (int sp = 0; sp < 2; ++sp)
                for (int c = 0; c < 3; ++c) {
                    const int d = 2 * c + sp;
                    Q[d + 6 * j] = psi[((int64_t)(2 * j + sp) * 3 + c) * vol + s];
                    A[d + 6 * j] = psi[((int64_t)(12 + 2 * j + sp) * 3 + c) * vol + s];
                }
        for (int kind = 0; kind < 4; ++kind) {
            const int na = (kind >= 2) ? 8 : 1;
            const int nk = (kind == 1 || kind == 3) ? 3 : 1;
            const double w = kind == 0 ? 1.0 : (kind == 1 ? 1.0 / 3.0 : (kind == 2 ? 6.0 / 8.0 : 6.0 / 24.0));
            cplx<double> tot = {0.0, 0.0};
            for (int a = 0; a < na; ++a)
                for (int k = 0; k < nk; ++k) {
                    // M_{(c,s),(c',s')} = X_{cc'} Y_{ss'}
                    cplx<double> M[36];
                    for (int p = 0; p < 6; ++p)
                        for (int q = 0; q < 6; ++q) {
                            const cplx<double> xv = kind >= 2 ? gen_elem(a, p / 2, q / 2) : cplx<double>{p / 2 == q / 2 ? 1.0 : 0.0, 0.0};
                            const cplx<double> yv = pauli_elem((kind == 1 || kind == 3) ? k + 1 : 0, p % 2, q % 2);
                            M[p + 6 * q] = xv * yv;
                        }
                    // N = M Q M ; tot += sum conj(A_pq) N_pq
                    for (int p = 0; p < 6; ++p)
                        for (int q = 0; q < 6; ++q) {
                            cplx<double> nv = {0.0, 0.0};
                            for (int rr = 0; rr < 6; ++rr) {
                                const cplx<double> mpr = M[p + 6 * rr];
                                if (mpr.re == 0.0 && mpr.im == 0.0) continue;
                                cplx<double> qm = {0.0, 0.0};
                                for (int t = 0; t < 6; ++t) qm = qm + Q[rr + 6 * t] * M[t + 6 * q];
                                nv = nv + mpr * qm;
                            }
                            tot = tot + cmul(A[p + 6 * q], nv);
                        }
                }
            // i * tot * w
            acc[2 * kind] += -tot.im * w;
            acc[2 * kind + 1] += tot.re * w;
        }
    }
    __shared__ double sh[8];
    for (int q = 0; q < 8; ++q) {
        const double v = block_sum(acc[q], sh);
        if (threadIdx.x == 0) partial[(int64_t)q * gridDim.x + blockIdx.x] = v;
    }
}

} // namespace nrq

using namespace nrq;

extern "C" int nrq_cn_init(double mass, const double c[8], double stepwidth, double ksp_tol)
{
    if (g_cn.initialised) return fail(NRQ_ERR_STATE, "nrqcd: already initialised");
    if (!lat().initialised) return fail(NRQ_ERR_STATE, "nrqcd: lattice module must be initialised first");
    if (!c) return fail(NRQ_ERR_ARG, "nrqcd: null Wilson coefficients");
    if (!(ksp_tol > 0) || !std::isfinite(mass) || mass == 0.0 || !std::isfinite(stepwidth))
        return fail(NRQ_ERR_ARG, "nrqcd: bad mass, step width or tolerance");
    if (int r = ensure_device()) return r;
    g_cn.mass = mass;
    g_cn.stepwidth = stepwidth;
    g_cn.solver = krylov_params();
    g_cn.solver.tol = ksp_tol;
    for (int k = 0; k < 4; ++k) g_cn.c[k] = std::complex<double>(c[2 * k], c[2 * k + 1]);
    g_cn.initialised = true;
    return NRQ_OK;
}

extern "C" int nrq_cn_finalize(void)
{
    if (!g_cn.initialised) return fail(NRQ_ERR_STATE, "nrqcd: not initialised");
    if (stream()) (void)hipStreamSynchronize(stream());
    krylov_release();
    g_cn.initialised = false;
    return NRQ_OK;
}

extern "C" int nrq_cn_set_solver(int kind, int restart, int64_t maxit)
{
    if (!g_cn.initialised) return fail(NRQ_ERR_STATE, "nrqcd: not initialised");
    if (kind < NRQ_SOLVER_GMRES_FP64 || kind > NRQ_SOLVER_FGMRES_CNN)
        return fail(NRQ_ERR_UNSUPPORTED, "nrqcd: solver kind %d not available", kind);
    if (restart < 1 || restart > 200 || maxit < 1) return fail(NRQ_ERR_ARG, "nrqcd: bad restart or maxit");
    static const krylov_kind kinds[] = {KRYLOV_GMRES_FP64, KRYLOV_GMRES_MIXED, KRYLOV_CHEB_FP64, KRYLOV_CHEB_MIXED,
                                        KRYLOV_FGMRES_CNN};
    g_cn.solver.kind = kinds[kind];
    // the fast solvers start from the first-order guess; plain GMRES and
    // FGMRES start from zero (PETSc's default), so that FGMRES differs from
    // GMRES only by the network
    g_cn.solver.extrapolate = kind != NRQ_SOLVER_GMRES_FP64 && kind != NRQ_SOLVER_FGMRES_CNN;
    g_cn.solver.restart = restart;
    g_cn.solver.maxit = maxit;
    return NRQ_OK;
}

#define NEED_HEAVY(h)                                                                          \
    do {                                                                                       \
        if (!(h)) return fail(NRQ_ERR_ARG, "heavy: null handle");                              \
        if (!(h)->psi) return fail(NRQ_ERR_STATE, "heavy: not allocated");                     \
        if ((h)->vol != lat().g.vol) return fail(NRQ_ERR_STATE, "heavy: lattice changed since allocation"); \
    } while (0)

static size_t heavy_bytes(int64_t v) { return (size_t)NFIELDS * 3 * (size_t)v * sizeof(cplx<double>); }

extern "C" int nrq_heavy_create(nrq_heavy **out)
{
    if (!lat().initialised) return fail(NRQ_ERR_STATE, "lattice: not initialised");
    if (!out) return fail(NRQ_ERR_ARG, "heavy: null output handle");
    if (int r = ensure_device()) return r;
    nrq_heavy *h = new nrq_heavy();
    h->vol = lat().g.vol;
    if (hipMalloc(&h->psi, heavy_bytes(h->vol)) != hipSuccess) {
        delete h;
        return fail(NRQ_ERR_ALLOC, "heavy: device allocation failed");
    }
    (void)hipMemsetAsync(h->psi, 0, heavy_bytes(h->vol), stream());
    h->live = 0;
    *out = h;
    return NRQ_OK;
}

extern "C" int nrq_heavy_destroy(nrq_heavy *h)
{
    if (!h) return NRQ_OK;
    if (stream()) (void)hipStreamSynchronize(stream());
    if (h->psi) (void)hipFree(h->psi);
    delete h;
    return NRQ_OK;
}

extern "C" int nrq_heavy_copy(nrq_heavy *dst, const nrq_heavy *src)
{
    NEED_HEAVY(src);
    if (!dst) return fail(NRQ_ERR_ARG, "heavy: null destination");
    if (dst == src) return NRQ_OK;
    if (!dst->psi || dst->vol != src->vol) {
        if (dst->psi) (void)hipFree(dst->psi);
        dst->psi = nullptr;
        dst->vol = src->vol;
        if (hipMalloc(&dst->psi, heavy_bytes(dst->vol)) != hipSuccess) return fail(NRQ_ERR_ALLOC, "heavy: allocation failed");
    }
    NRQ_HIP(hipMemcpyAsync(dst->psi, src->psi, heavy_bytes(src->vol), hipMemcpyDeviceToDevice, stream()));
    dst->live = src->live;
    return NRQ_OK;
}

extern "C" int nrq_heavy_init_point(nrq_heavy *h, int64_t iq, int64_t ia)
{
    NEED_HEAVY(h);
    if (iq < 1 || iq > h->vol || ia < 1 || ia > h->vol) return fail(NRQ_ERR_ARG, "heavy: lattice index out of range");
    NRQ_HIP(hipMemsetAsync(h->psi, 0, heavy_bytes(h->vol), stream()));
    const double v = 1.0 / std::sqrt(6.0);
    hipLaunchKernelGGL(k_set_point, dim3(1), dim3(64), 0, stream(), (cplx<double> *)h->psi, h->vol, 0, iq - 1, v);
    hipLaunchKernelGGL(k_set_point, dim3(1), dim3(64), 0, stream(), (cplx<double> *)h->psi, h->vol, 12, ia - 1, v);
    NRQ_HIP(hipGetLastError());
    // fields (j, spin) with j - spin even carry the source, for both propagators
    h->live = 0;
    for (int t = 0; t < 12; ++t)
        if ((t / 2 - t % 2) % 2 == 0) h->live |= (1u << t) | (1u << (12 + t));
    return NRQ_OK;
}

extern "C" int nrq_heavy_init_point_dof(nrq_heavy *h, int spin_q, int colour_q, int64_t iq, int spin_a, int colour_a,
                                        int64_t ia)
{
    NEED_HEAVY(h);
    if (iq < 1 || iq > h->vol || ia < 1 || ia > h->vol) return fail(NRQ_ERR_ARG, "heavy: lattice index out of range");
    if (spin_q < 1 || spin_q > 2 || spin_a < 1 || spin_a > 2 || colour_q < 1 || colour_q > 3 || colour_a < 1 || colour_a > 3)
        return fail(NRQ_ERR_ARG, "heavy: spin 1..2, colour 1..3");
    NRQ_HIP(hipMemsetAsync(h->psi, 0, heavy_bytes(h->vol), stream()));
    // nrqcd.f90:437-444: index = colour + 3 (spin - 1) (1-based), used as both
    // row and column of the 6x6 propagator.  In the C2CS row ordering d =
    // 2(c-1) + s, that dof index D (0-based) is colour D/2, spin D%2.
    const int dofs[2] = {colour_q + 3 * (spin_q - 1) - 1, colour_a + 3 * (spin_a - 1) - 1};
    const int64_t sites[2] = {iq - 1, ia - 1};
    h->live = 0;
    for (int p = 0; p < 2; ++p) {
        const int D = dofs[p];
        const int col = D, c = D / 2, sp = D % 2;
        const int field = 12 * p + 2 * col + sp;
        h->live |= 1u << field;
        const cplx<double> one = {1.0, 0.0};
        NRQ_HIP(hipMemcpyAsync((cplx<double> *)h->psi + ((int64_t)field * 3 + c) * h->vol + sites[p], &one, sizeof(one),
                               hipMemcpyHostToDevice, stream()));
    }
    NRQ_HIP(hipStreamSynchronize(stream()));
    return NRQ_OK;
}

extern "C" int nrq_heavy_upload(nrq_heavy *h, const double *quark, const double *antiq)
{
    NEED_HEAVY(h);
    const int64_t v = h->vol;
    std::vector<cplx<double>> buf((size_t)12 * 3 * (size_t)v);
    const double *src[2] = {quark, antiq};
    for (int p = 0; p < 2; ++p) {
        if (!src[p]) continue;
        unsigned live = 0;
        for (int64_t s = 0; s < v; ++s)
            for (int j = 0; j < 6; ++j)
                for (int d = 0; d < 6; ++d) {
                    const int c = d / 2, sp = d % 2;
                    const double *z = src[p] + 2 * (d + 6 * j + 36 * s);
                    buf[((size_t)(2 * j + sp) * 3 + c) * (size_t)v + (size_t)s] = {z[0], z[1]};
                    if (z[0] != 0.0 || z[1] != 0.0) live |= 1u << (12 * p + 2 * j + sp);
                }
        h->live = (h->live & ~(0xfffu << (12 * p))) | live;
        NRQ_HIP(hipMemcpyAsync((cplx<double> *)h->psi + (size_t)12 * p * 3 * (size_t)v, buf.data(), buf.size() * sizeof(cplx<double>),
                               hipMemcpyHostToDevice, stream()));
        NRQ_HIP(hipStreamSynchronize(stream()));
    }
    return NRQ_OK;
}

extern "C" int nrq_heavy_download(const nrq_heavy *h, double *quark, double *antiq)
{
    NEED_HEAVY(h);
    const int64_t v = h->vol;
    std::vector<cplx<double>> buf((size_t)12 * 3 * (size_t)v);
    double *dst[2] = {quark, antiq};
    for (int p = 0; p < 2; ++p) {
        if (!dst[p]) continue;
        NRQ_HIP(hipMemcpyAsync(buf.data(), (const cplx<double> *)h->psi + (size_t)12 * p * 3 * (size_t)v, buf.size() * sizeof(cplx<double>),
                               hipMemcpyDeviceToHost, stream()));
        NRQ_HIP(hipStreamSynchronize(stream()));
        // C2CS structure: rows of different spin than the field's spin are zero
        for (int64_t s = 0; s < v; ++s)
            for (int j = 0; j < 6; ++j)
                for (int d = 0; d < 6; ++d) {
                    const int c = d / 2, sp = d % 2;
                    const cplx<double> z = buf[((size_t)(2 * j + sp) * 3 + c) * (size_t)v + (size_t)s];
                    dst[p][2 * (d + 6 * j + 36 * s)] = z.re;
                    dst[p][2 * (d + 6 * j + 36 * s) + 1] = z.im;
                }
    }
    return NRQ_OK;
}

// One CN step of the live fields of n heavy fields on one gauge
// configuration as a single batched solve; heavy field k has mass[k].
static int heavy_update_n(nrq_heavy *const *hs, const double *mass, int n, const nrq_gauge *gf, const double c[8],
                          double stepwidth, nrq_solve_stats *stats)
{
    if (!g_cn.initialised) return fail(NRQ_ERR_STATE, "nrqcd: not initialised");
    if (!hs || !mass || n < 1) return fail(NRQ_ERR_ARG, "heavy: no fields");
    if (!gf || !gf->u[0]) return fail(NRQ_ERR_ARG, "heavy: null gauge configuration");
    if (!c) return fail(NRQ_ERR_ARG, "heavy: null Wilson coefficients");
    if (!std::isfinite(stepwidth)) return fail(NRQ_ERR_ARG, "heavy: bad step width");
    for (int k = 1; k < 4; ++k)
        if (c[2 * k] != 0.0 || c[2 * k + 1] != 0.0)
            return fail(NRQ_ERR_UNSUPPORTED, "heavy: only the kinetic term c1 is implemented (c%d != 0)", k + 1);
    const double a0 = lat().g.a[0];
    const std::complex<double> c1(c[0], c[1]);
    const std::complex<double> mi(0.0, -(stepwidth / 2.0 * a0));
    // The live fields are the right-hand sides (a field that is exactly zero
    // stays zero, as KSPSolve returns x = 0 for b = 0).  rhs = S psi with
    // S = 1 + alpha K (SystMat_herm before the in-place Hermitian
    // transpose), then S^+ psi' = rhs with S^+ = 1 + conj(alpha) K.
    cplx<double> *fields[NSLOTS];
    cplx<double> alpha[NSLOTS], beta[NSLOTS];
    int R = 0;
    for (int k = 0; k < n; ++k) {
        const nrq_heavy *h = hs[k];
        NEED_HEAVY(h);
        if (gf->vol != h->vol) return fail(NRQ_ERR_ARG, "heavy: gauge and heavy field lattices differ");
        if (!std::isfinite(mass[k]) || mass[k] == 0.0) return fail(NRQ_ERR_ARG, "heavy: bad mass %g", mass[k]);
        for (int j = 0; j < k; ++j)
            if (hs[j] == h) return fail(NRQ_ERR_ARG, "heavy: field %d listed twice", k);
        const std::complex<double> alpha_q = mi * c1 / mass[k];
        const std::complex<double> alpha_a = mi * std::conj(c1) / (-mass[k]);
        for (int f = 0; f < NFIELDS; ++f) {
            if (!(h->live >> f & 1u)) continue;
            if (R == NSLOTS) return fail(NRQ_ERR_ARG, "heavy: more than %d live fields in one batch", NSLOTS);
            const std::complex<double> a = f < 12 ? alpha_q : alpha_a;
            fields[R] = (cplx<double> *)h->psi + (int64_t)f * 3 * h->vol;
            alpha[R] = {a.real(), a.imag()};
            beta[R] = {a.real(), -a.imag()};
            ++R;
        }
    }
    nrq_solve_stats st = {0, 0, 0, 0.0};
    st.solves = R;
    if (int r = krylov_reserve(hs[0]->vol, g_cn.solver.restart, R)) return r;
    const int status = krylov_cn_step(gf, fields, R, alpha, beta, g_cn.solver, &st);
    if (status != NRQ_OK && status != NRQ_ERR_NOCONV) return status;
    NRQ_HIP(hipStreamSynchronize(stream()));
    if (stats) *stats = st;
    return status;
}

extern "C" int nrq_heavy_update(nrq_heavy *h, const nrq_gauge *gf, double mass, const double c[8], double stepwidth,
                                nrq_solve_stats *stats)
{
    return heavy_update_n(&h, &mass, 1, gf, c, stepwidth, stats);
}

extern "C" int nrq_heavy_update_batch(nrq_heavy *const *h, const double *mass, int n, const nrq_gauge *gf,
                                      const double c[8], double stepwidth, nrq_solve_stats *stats)
{
    return heavy_update_n(h, mass, n, gf, c, stepwidth, stats);
}

extern "C" int nrq_heavy_norms(const nrq_heavy *h, double *nq, double *na)
{
    NEED_HEAVY(h);
    if (!nq || !na) return fail(NRQ_ERR_ARG, "null argument");
    if (int r = krylov_reserve(h->vol, g_cn.initialised ? g_cn.solver.restart : 30, 0)) return r;
    double n2[NFIELDS];
    if (int e = krylov_norms2((const cplx<double> *)h->psi, NFIELDS, n2)) return e;
    double q = 0, a = 0;
    for (int r = 0; r < 12; ++r) q += n2[r];
    for (int r = 12; r < 24; ++r) a += n2[r];
    *nq = std::sqrt(q);
    *na = std::sqrt(a);
    return NRQ_OK;
}

extern "C" int nrq_heavy_correlators(const nrq_heavy *h, double out[8])
{
    NEED_HEAVY(h);
    if (!out) return fail(NRQ_ERR_ARG, "null argument");
    if (int r = krylov_reserve(h->vol, g_cn.initialised ? g_cn.solver.restart : 30, 0)) return r;
    const int nb = 256;
    double *part = krylov_reduce_scratch();
    double *res = part + 8 * nb;
    hipLaunchKernelGGL(k_corr, dim3(nb), dim3(REDUCE_THREADS), 0, stream(), (const cplx<double> *)h->psi, h->vol, part);
    hipLaunchKernelGGL(k_fold8, dim3(1), dim3(64), 0, stream(), part, nb, res);
    NRQ_HIP(hipGetLastError());
    double *hostbuf = krylov_host_scratch();
    NRQ_HIP(hipMemcpyAsync(hostbuf, res, sizeof(double) * 8, hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    memcpy(out, hostbuf, sizeof(double) * 8);
    return NRQ_OK;
}
