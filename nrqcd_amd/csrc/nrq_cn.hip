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

// Zero-momentum meson correlators (nrqcd.f90:1484-1661, and the _gconjg
// variants :1728-1905) as traces.  Each is sum_x i Tr(P M Q M) w over
// vertices M = X (x) Y, colour-spin index d = 2c + s, with P = A^+ (plain) or
// P = conj(A) (_gconjg).  The vertex sums collapse through
//   sum_a T^a_ij T^a_kl = (delta_il delta_jk - delta_ij delta_kl / 3) / 2,
//   sum_k sigma^k_ij sigma^k_kl = 2 delta_il delta_jk - delta_ij delta_kl,
// which hold for the reference's tables (su3.f90:97-131, su2.f90:68-88):
// they differ from the textbook matrices by signs only, and every vertex
// enters twice.  What remains are four traces per site,
//   t0 = Tr(P Q),  tX = Tr P Tr Q,
//   tC = Tr(Ps Qs), Ps_ss' = sum_c P_(c,s),(c,s')   (colour traced),
//   tS = Tr(Pc Qc), Pc_cc' = sum_s P_(c,s),(c',s)   (spin traced),
// and O1(1S0) = i t0, O1(3S1) = i (2 tS - t0)/3, O8(1S0) = i 3/8 (tC - t0/3),
// O8(3S1) = i/8 (2 tX - tC - (2 tS - t0)/3); the host forms these from the
// lattice sums.  The same pass sums |Q|^2 and |A|^2 for the norms
// (nrqcd.f90:547-611).  Per site the kernel reads the 72 complex numbers of Q and A
// once and produces both sets: with As, Ac the partial traces of A,
//   plain:   t0 = sum_e conj(A_e) Q_e,      tC = sum_e conj(As_e) Qs_e, ...
//   _gconjg: t0 = sum_xy conj(A_xy) Q_yx,   tC = sum_ss' conj(As_ss') Qs_s's, ...
// Sums out[v * gridDim.x + block]: v = 0..13 = re/im of t0, tX, tC, tS (plain),
// t0, tC, tS (_gconjg), tX being the same for both; v = 14, 15: |Q|^2, |A|^2.
constexpr int CORR_SUMS = 16;

// Accumulates one propagator entry (row r, column c) of Q and A into the
// partial traces, traces and squared norms.
struct corr_acc {
    cplx<double> qs[4], as[4], qc[9], ac[9], trq, tra;
    double nq, na;
    __device__ __forceinline__ void entry(int r, int c, cplx<double> q, cplx<double> a)
    {
        nq += q.re * q.re + q.im * q.im;
        na += a.re * a.re + a.im * a.im;
        if ((r >> 1) == (c >> 1)) { // same colour: spin block of the colour trace
            qs[(r & 1) + 2 * (c & 1)] = qs[(r & 1) + 2 * (c & 1)] + q;
            as[(r & 1) + 2 * (c & 1)] = as[(r & 1) + 2 * (c & 1)] + a;
        }
        if ((r & 1) == (c & 1)) { // same spin: colour block of the spin trace
            qc[(r >> 1) + 3 * (c >> 1)] = qc[(r >> 1) + 3 * (c >> 1)] + q;
            ac[(r >> 1) + 3 * (c >> 1)] = ac[(r >> 1) + 3 * (c >> 1)] + a;
        }
        if (r == c) {
            trq = trq + q;
            tra = tra + a;
        }
    }
};

__global__ __launch_bounds__(REDUCE_THREADS) void k_corr_traces(const cplx<double> *__restrict__ psi, int64_t vol,
                                                                double *__restrict__ partial)
{
    cplx<double> t0 = {0, 0}, tx = {0, 0}, tc = {0, 0}, ts = {0, 0}, g0 = {0, 0}, gc = {0, 0}, gs = {0, 0};
    double nq = 0, na = 0;
    // entry (d, j) of the quark (field 2j + (d & 1), colour plane d >> 1) and
    // of the antiquark (fields 12 + ...)
    auto at = [&](int f0, int d, int j, int64_t s) { return psi[((int64_t)(f0 + 2 * j + (d & 1)) * 3 + (d >> 1)) * vol + s]; };
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < vol; s += (int64_t)gridDim.x * blockDim.x) {
        corr_acc w;
#pragma unroll
        for (int e = 0; e < 4; ++e) w.qs[e] = w.as[e] = {0, 0};
#pragma unroll
        for (int e = 0; e < 9; ++e) w.qc[e] = w.ac[e] = {0, 0};
        w.trq = w.tra = {0, 0};
        w.nq = w.na = 0;
        // entries in transposed pairs (d, j), (j, d): t0 pairs equal
        // positions, the _gconjg t0 pairs transposed ones
#pragma unroll
        for (int d = 0; d < 6; ++d)
#pragma unroll
            for (int j = d; j < 6; ++j) {
                const cplx<double> q1 = at(0, d, j, s), a1 = at(12, d, j, s);
                w.entry(d, j, q1, a1);
                t0 = t0 + cmul(a1, q1);
                if (j == d) {
                    g0 = g0 + cmul(a1, q1);
                } else {
                    const cplx<double> q2 = at(0, j, d, s), a2 = at(12, j, d, s);
                    w.entry(j, d, q2, a2);
                    t0 = t0 + cmul(a2, q2);
                    g0 = g0 + cmul(a1, q2) + cmul(a2, q1);
                }
            }
        nq += w.nq;
        na += w.na;
        tx = tx + cmul(w.tra, w.trq);
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                tc = tc + cmul(w.as[r + 2 * c], w.qs[r + 2 * c]);
                gc = gc + cmul(w.as[c + 2 * r], w.qs[r + 2 * c]);
            }
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                ts = ts + cmul(w.ac[r + 3 * c], w.qc[r + 3 * c]);
                gs = gs + cmul(w.ac[c + 3 * r], w.qc[r + 3 * c]);
            }
    }
    const double v[CORR_SUMS] = {t0.re, t0.im, tx.re, tx.im, tc.re, tc.im, ts.re, ts.im,
                                 g0.re, g0.im, gc.re, gc.im, gs.re, gs.im, nq, na};
    __shared__ double sh[REDUCE_THREADS / 64];
#pragma unroll
    for (int q = 0; q < CORR_SUMS; ++q) {
        const double b = block_sum(v[q], sh);
        if (threadIdx.x == 0) partial[(int64_t)q * gridDim.x + blockIdx.x] = b;
    }
}

// out[q] = sum of the nblk partials of value q; one block per value, fixed
// order (thread t sums t, t + 256, ..., then the block sum).
__global__ __launch_bounds__(REDUCE_THREADS) void k_fold(const double *__restrict__ partial, int nblk,
                                                         double *__restrict__ out)
{
    const int q = blockIdx.x;
    double s = 0;
    for (int b = threadIdx.x; b < nblk; b += REDUCE_THREADS) s += partial[(int64_t)q * nblk + b];
    __shared__ double sh[REDUCE_THREADS / 64];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[q] = s;
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
    if (kind < NRQ_SOLVER_GMRES_FP64 || kind > NRQ_SOLVER_FGMRES_CNN_MIXED)
        return fail(NRQ_ERR_UNSUPPORTED, "nrqcd: solver kind %d not available", kind);
    if (restart < 1 || restart > 200 || maxit < 1) return fail(NRQ_ERR_ARG, "nrqcd: bad restart or maxit");
    static const krylov_kind kinds[] = {KRYLOV_GMRES_FP64, KRYLOV_GMRES_MIXED, KRYLOV_CHEB_FP64, KRYLOV_CHEB_MIXED,
                                        KRYLOV_FGMRES_CNN, KRYLOV_FGMRES_CNN_MIXED};
    g_cn.solver.kind = kinds[kind];
    // the fast solvers start from the first-order guess; plain GMRES and
    // FGMRES start from zero (PETSc's default), so that FGMRES differs from
    // GMRES only by the network (and the mixed FGMRES from the mixed GMRES)
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
    ++h->version;
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
    ++dst->version;
    dst->obs_version = ~(uint64_t)0;
    if (src->obs_version == src->version) { // the copy has the same observables
        std::memcpy(dst->obs, src->obs, sizeof(dst->obs));
        dst->obs_version = dst->version;
    }
    return NRQ_OK;
}

extern "C" int nrq_heavy_init_point(nrq_heavy *h, int64_t iq, int64_t ia)
{
    NEED_HEAVY(h);
    if (iq < 1 || iq > h->vol || ia < 1 || ia > h->vol) return fail(NRQ_ERR_ARG, "heavy: lattice index out of range");
    ++h->version;
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
    ++h->version;
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
        ++h->version;
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
    bool full = false; // any of c2..c4: the full Hamiltonian (nrq_hamiltonian.hip)
    for (int k = 1; k < 4; ++k) full = full || c[2 * k] != 0.0 || c[2 * k + 1] != 0.0;
    const double a0 = lat().g.a[0];
    const std::complex<double> c1(c[0], c[1]);
    const std::complex<double> mi(0.0, -(stepwidth / 2.0 * a0));
    // The live fields are the right-hand sides (a field that is exactly zero
    // stays zero, as KSPSolve returns x = 0 for b = 0).  rhs = S psi with
    // S = 1 + alpha K (SystMat_herm before the in-place Hermitian
    // transpose), then S^+ psi' = rhs with S^+ = 1 + conj(alpha) K.
    if (full) {
        for (int k = 0; k < n; ++k) {
            const nrq_heavy *h = hs[k];
            NEED_HEAVY(h);
            if (gf->vol != h->vol) return fail(NRQ_ERR_ARG, "heavy: gauge and heavy field lattices differ");
            if (!std::isfinite(mass[k]) || mass[k] == 0.0) return fail(NRQ_ERR_ARG, "heavy: bad mass %g", mass[k]);
            for (int j = 0; j < k; ++j)
                if (hs[j] == h) return fail(NRQ_ERR_ARG, "heavy: field %d listed twice", k);
        }
        for (int k = 0; k < n; ++k) ++hs[k]->version; // the solve overwrites the live columns
        nrq_solve_stats st = {0, 0, 0, 0.0};
        const int status = ham_cn_update(hs, mass, n, gf, c, stepwidth, g_cn.solver.restart, g_cn.solver.maxit,
                                         g_cn.solver.tol, &st);
        if (status != NRQ_OK && status != NRQ_ERR_NOCONV) return status;
        NRQ_HIP(hipStreamSynchronize(stream()));
        if (stats) *stats = st;
        return status;
    }
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
    for (int k = 0; k < n; ++k) ++hs[k]->version; // the solve overwrites the live fields
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

// Blocks of one k_corr_traces pass: one site per thread, at most
// CORR_BLOCKS_MAX blocks (the kernel strides over larger lattices).
constexpr int CORR_BLOCKS_MAX = 4096;
static int corr_blocks(int64_t vol)
{
    return (int)std::min<int64_t>(CORR_BLOCKS_MAX, std::max<int64_t>(1, (vol + REDUCE_THREADS - 1) / REDUCE_THREADS));
}

// Brings h->obs up to date with the field (one k_corr_traces pass when the
// field changed since the last evaluation).
static int heavy_observables(const nrq_heavy *hc)
{
    NEED_HEAVY(hc);
    nrq_heavy *h = const_cast<nrq_heavy *>(hc); // the cache is not part of the field's value
    if (h->obs_version == h->version) return NRQ_OK;
    if (int r = krylov_reserve(h->vol, g_cn.initialised ? g_cn.solver.restart : 30, 0)) return r;
    // one site per thread up to 4096 blocks (the partials, 16 x 4096 doubles,
    // and the 16 sums fit the reduction scratch)
    static_assert((size_t)CORR_SUMS * CORR_BLOCKS_MAX + CORR_SUMS <= KRYLOV_REDUCE_DOUBLES, "correlator partials");
    const int nb = corr_blocks(h->vol);
    double *part = krylov_reduce_scratch();
    double *res = part + (size_t)CORR_SUMS * nb;
    hipLaunchKernelGGL(k_corr_traces, dim3(nb), dim3(REDUCE_THREADS), 0, stream(), (const cplx<double> *)h->psi,
                       h->vol, part);
    hipLaunchKernelGGL(k_fold, dim3(CORR_SUMS), dim3(REDUCE_THREADS), 0, stream(), part, nb, res);
    NRQ_HIP(hipGetLastError());
    double *hostbuf = krylov_host_scratch();
    NRQ_HIP(hipMemcpyAsync(hostbuf, res, sizeof(double) * CORR_SUMS, hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    std::memcpy(h->obs, hostbuf, sizeof(double) * CORR_SUMS);
    h->obs_version = h->version;
    return NRQ_OK;
}

extern "C" int nrq_heavy_norms(const nrq_heavy *h, double *nq, double *na)
{
    if (!nq || !na) return fail(NRQ_ERR_ARG, "null argument");
    if (int r = heavy_observables(h)) return r;
    *nq = std::sqrt(h->obs[14]);
    *na = std::sqrt(h->obs[15]);
    return NRQ_OK;
}

// The plain (gconjg = false) or _gconjg correlators from the CORR_SUMS trace
// sums t.
static void correlators_from_sums(const double *t, bool gconjg, double out[8])
{
    using zc = std::complex<double>;
    const int o = gconjg ? 8 : 0;
    const zc t0(t[o], t[o + 1]), tx(t[2], t[3]);
    const zc tc(t[gconjg ? 10 : 4], t[gconjg ? 11 : 5]);
    const zc ts(t[gconjg ? 12 : 6], t[gconjg ? 13 : 7]);
    const zc i(0, 1);
    const zc c[4] = {i * t0, i * (2.0 * ts - t0) / 3.0, i * 0.375 * (tc - t0 / 3.0),
                     i * 0.125 * (2.0 * tx - tc - (2.0 * ts - t0) / 3.0)};
    for (int k = 0; k < 4; ++k) {
        out[2 * k] = c[k].real();
        out[2 * k + 1] = c[k].imag();
    }
}

static int heavy_correlators(const nrq_heavy *h, bool gconjg, double out[8])
{
    if (!out) return fail(NRQ_ERR_ARG, "null argument");
    if (int r = heavy_observables(h)) return r;
    correlators_from_sums(h->obs, gconjg, out);
    return NRQ_OK;
}

extern "C" int nrq_heavy_correlators(const nrq_heavy *h, double out[8]) { return heavy_correlators(h, false, out); }

extern "C" int nrq_heavy_correlators_gconjg(const nrq_heavy *h, double out[8])
{
    return heavy_correlators(h, true, out);
}

// ---------------------------------------------------------------------------
// Allocation-free correlators.  The partial sums live in `scratch`, a
// fixed-size device array of this file: the partials of CORR_BLOCKS_MAX
// blocks and the CORR_SUMS folded sums (524,416 bytes), enough for any
// lattice because k_corr_traces strides over the sites of larger ones.  Each
// GPU holds its own copy until the process ends; its address is looked up
// once per device (hipGetSymbolAddress, with that device current) and kept,
// and only that address is handed to kernels.  Every block writes its
// partials and the fold writes every sum, so nothing needs clearing between
// calls; calls that overlap on one GPU (other host threads, other streams)
// share the array and race.
// ---------------------------------------------------------------------------
namespace nrq {
constexpr size_t CORR_SCRATCH_DOUBLES = (size_t)CORR_SUMS * (CORR_BLOCKS_MAX + 1);
static __device__ __align__(256) double scratch[CORR_SCRATCH_DOUBLES];
constexpr int SCRATCH_MAX_DEVICES = 64;
static double *g_scratch_dev[SCRATCH_MAX_DEVICES]; // device address of `scratch`, per device

static int corr_scratch(double **out)
{
    int dev = -1;
    NRQ_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES)
        return fail(NRQ_ERR_ARG, "heavy: device %d beyond the %d devices the correlator scratch serves", dev,
                    SCRATCH_MAX_DEVICES);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        NRQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(NRQ_ERR_HIP, "heavy: no device address for the correlator scratch on device %d", dev);
        g_scratch_dev[dev] = (double *)p;
    }
    *out = g_scratch_dev[dev];
    return NRQ_OK;
}
} // namespace nrq

extern "C" int nrq_heavy_correlators_noalloc(const nrq_heavy *h, double out[8])
{
    if (!out) return fail(NRQ_ERR_ARG, "null argument");
    NEED_HEAVY(h);
    double *part = nullptr;
    if (int r = corr_scratch(&part)) return r;
    const int nb = corr_blocks(h->vol);
    if ((size_t)CORR_SUMS * ((size_t)nb + 1) > CORR_SCRATCH_DOUBLES)
        return fail(NRQ_ERR_ARG, "heavy: %d partial blocks exceed the correlator scratch of %d", nb, CORR_BLOCKS_MAX);
    double *res = part + (size_t)CORR_SUMS * nb;
    hipLaunchKernelGGL(k_corr_traces, dim3(nb), dim3(REDUCE_THREADS), 0, stream(), (const cplx<double> *)h->psi,
                       h->vol, part);
    hipLaunchKernelGGL(k_fold, dim3(CORR_SUMS), dim3(REDUCE_THREADS), 0, stream(), part, nb, res);
    NRQ_HIP(hipGetLastError());
    double sums[CORR_SUMS];
    NRQ_HIP(hipMemcpyAsync(sums, res, sizeof(sums), hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    correlators_from_sums(sums, false, out);
    return NRQ_OK;
}
