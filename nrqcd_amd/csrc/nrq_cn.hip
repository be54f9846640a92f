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
#include "nrq_internal.h"
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

namespace nrq {

const void *gauge_links(const nrq_gauge *h);

constexpr int NFIELDS = 24;
constexpr int RCHUNK = 4;       // fields per thread in the stencil (link reuse)
constexpr int VBLOCKS = 256;    // blocks per field in the vector kernels
constexpr int VTHREADS = 256;

struct cn_config {
    bool initialised = false;
    double mass = 1, stepwidth = 1, tol = 1e-10;
    std::complex<double> c[4];
    int kind = NRQ_SOLVER_GMRES_FP64;
    int restart = 30;
    int64_t maxit = 10000;
};
static cn_config g_cn;

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// out_r = in_r + beta_r K in_r for the fields r in [r0, r0+RCHUNK) of a
// chunk; links loaded once per (site, direction) and reused across fields.
template <typename RL>
__global__ __launch_bounds__(256) void k_stencil(const cplx<RL> *__restrict__ u, const cplx<double> *__restrict__ in,
                                                 cplx<double> *__restrict__ out, const cplx<double> *__restrict__ beta,
                                                 const int *__restrict__ active, int nfields, geom g)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r0 = blockIdx.y * RCHUNK;
    if (s >= g.vol) return;
    const int64_t fs = 3 * g.vol; // field stride
    cplx<double> acc[RCHUNK][3];
    bool on[RCHUNK];
#pragma unroll
    for (int q = 0; q < RCHUNK; ++q) {
        on[q] = (r0 + q < nfields) && active[r0 + q];
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[q][c] = {0.0, 0.0};
    }
    for (int i = 0; i < 3; ++i) {
        const double ia2 = 1.0 / (g.a[i + 1] * g.a[i + 1]);
        const int64_t sp = neib_up(g, s, i), sm = neib_dn(g, s, i);
        cplx<double> up[9], um[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            const cplx<RL> a = u[((int64_t)i * 9 + e) * g.vol + s];
            const cplx<RL> b = u[((int64_t)i * 9 + e) * g.vol + sm];
            up[e] = {(double)a.re, (double)a.im};
            um[e] = {(double)b.re, (double)b.im};
        }
#pragma unroll
        for (int q = 0; q < RCHUNK; ++q) {
            if (!on[q]) continue;
            const cplx<double> *f = in + (int64_t)(r0 + q) * fs;
            cplx<double> vp[3], vm[3], v0[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                vp[c] = f[(int64_t)c * g.vol + sp];
                vm[c] = f[(int64_t)c * g.vol + sm];
                v0[c] = f[(int64_t)c * g.vol + s];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                // U_i(x) psi(x+i) + U_i(x-i)^+ psi(x-i)
                cplx<double> hop = up[3 * c] * vp[0] + up[3 * c + 1] * vp[1] + up[3 * c + 2] * vp[2];
                hop = hop + cmul(um[c], vm[0]) + cmul(um[3 + c], vm[1]) + cmul(um[6 + c], vm[2]);
                acc[q][c] = acc[q][c] + ia2 * (v0[c] - 0.5 * hop);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RCHUNK; ++q) {
        if (!on[q]) continue;
        const int r = r0 + q;
        const cplx<double> b = beta[r];
        const cplx<double> *f = in + (int64_t)r * fs;
        cplx<double> *o = out + (int64_t)r * fs;
#pragma unroll
        for (int c = 0; c < 3; ++c) o[(int64_t)c * g.vol + s] = f[(int64_t)c * g.vol + s] + b * acc[q][c];
    }
}

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Block-level sum of a double across 256 threads; result valid in thread 0.
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

// partial[r][i][blk] = sum over this block's elements of conj(V_i,r) w_r,
// i = 0..nv-1.  Basis V[i][r][n], w[r][n], n = 3 vol.
__global__ __launch_bounds__(VTHREADS) void k_mdot(const cplx<double> *__restrict__ basis, int nv,
                                                   const cplx<double> *__restrict__ w, int64_t n, int nfields,
                                                   const int *__restrict__ active, cplx<double> *__restrict__ partial)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    __shared__ double sh[8];
    const cplx<double> *wr = w + (int64_t)r * n;
    for (int i = 0; i < nv; ++i) {
        const cplx<double> *vr = basis + ((int64_t)i * nfields + r) * n;
        double re = 0, im = 0;
        for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
            const cplx<double> a = vr[k], b = wr[k];
            re += a.re * b.re + a.im * b.im;
            im += a.re * b.im - a.im * b.re;
        }
        re = block_sum(re, sh);
        im = block_sum(im, sh);
        if (threadIdx.x == 0) partial[((int64_t)r * nv + i) * gridDim.x + blockIdx.x] = {re, im};
    }
}

// out[r*count + i] = sum_blk partial[(r*count + i)*nblk + blk] (fixed order)
__global__ void k_fold_c(const cplx<double> *__restrict__ partial, int nblk, int count, int nfields,
                         const int *__restrict__ active, cplx<double> *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * nfields) return;
    const int r = t / count;
    if (!active[r]) return;
    cplx<double> s = {0.0, 0.0};
    for (int b = 0; b < nblk; ++b) s = s + partial[(int64_t)t * nblk + b];
    out[t] = s;
}

// CGS update w_r -= sum_i h[r][i] V_i,r and partial ||w_r||^2.
// h is the device copy of the projections just folded by k_fold_c.
__global__ __launch_bounds__(VTHREADS) void k_cgs_update(const cplx<double> *__restrict__ basis, int nv,
                                                         cplx<double> *__restrict__ w, int64_t n, int nfields,
                                                         const int *__restrict__ active,
                                                         const cplx<double> *__restrict__ h, int hstride,
                                                         double *__restrict__ partial)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    __shared__ double sh[8];
    cplx<double> *wr = w + (int64_t)r * n;
    double nrm = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        cplx<double> a = wr[k];
        for (int i = 0; i < nv; ++i) a = a - h[(int64_t)r * hstride + i] * basis[((int64_t)i * nfields + r) * n + k];
        wr[k] = a;
        nrm += a.re * a.re + a.im * a.im;
    }
    nrm = block_sum(nrm, sh);
    if (threadIdx.x == 0) partial[(int64_t)r * gridDim.x + blockIdx.x] = nrm;
}

__global__ void k_fold_r(const double *__restrict__ partial, int nblk, int nfields, const int *__restrict__ active,
                         double *__restrict__ out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nfields || !active[r]) return;
    double s = 0;
    for (int b = 0; b < nblk; ++b) s += partial[(int64_t)r * nblk + b];
    out[r] = s;
}

// dst_r = src_r / sqrt(nrm2[r]) (zero norm leaves zeros)
__global__ void k_scale(const cplx<double> *__restrict__ src, cplx<double> *__restrict__ dst, int64_t n,
                        const double *__restrict__ nrm2, const int *__restrict__ active)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    const double nn = nrm2[r];
    const double inv = nn > 0 ? 1.0 / sqrt(nn) : 0.0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        dst[(int64_t)r * n + k] = inv * src[(int64_t)r * n + k];
}

// x_r += sum_i y[r][i] V_i,r
__global__ void k_xupdate(cplx<double> *__restrict__ x, const cplx<double> *__restrict__ basis, int nfields,
                          int64_t n, const cplx<double> *__restrict__ y, int ystride, const int *__restrict__ nused,
                          const int *__restrict__ active)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    const int nv = nused[r];
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        cplx<double> a = x[(int64_t)r * n + k];
        for (int i = 0; i < nv; ++i) a = a + y[(int64_t)r * ystride + i] * basis[((int64_t)i * nfields + r) * n + k];
        x[(int64_t)r * n + k] = a;
    }
}

// res_r = b_r - t_r and partial ||res_r||^2
__global__ __launch_bounds__(VTHREADS) void k_residual(const cplx<double> *__restrict__ b, const cplx<double> *__restrict__ t,
                                                       cplx<double> *__restrict__ res, int64_t n,
                                                       const int *__restrict__ active, double *__restrict__ partial)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    __shared__ double sh[8];
    double nrm = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const cplx<double> a = b[(int64_t)r * n + k] - t[(int64_t)r * n + k];
        res[(int64_t)r * n + k] = a;
        nrm += a.re * a.re + a.im * a.im;
    }
    nrm = block_sum(nrm, sh);
    if (threadIdx.x == 0) partial[(int64_t)r * gridDim.x + blockIdx.x] = nrm;
}

// partial ||x_r||^2
__global__ __launch_bounds__(VTHREADS) void k_norm2(const cplx<double> *__restrict__ x, int64_t n,
                                                    const int *__restrict__ active, double *__restrict__ partial)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    __shared__ double sh[8];
    double nrm = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const cplx<double> a = x[(int64_t)r * n + k];
        nrm += a.re * a.re + a.im * a.im;
    }
    nrm = block_sum(nrm, sh);
    if (threadIdx.x == 0) partial[(int64_t)r * gridDim.x + blockIdx.x] = nrm;
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

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct cn_workspace {
    int64_t vol = 0;
    int m = 0;
    cplx<double> *basis = nullptr; // (m+1) x NFIELDS x 3V
    cplx<double> *b = nullptr, *x = nullptr, *t = nullptr, *res = nullptr; // NFIELDS x 3V each
    cplx<double> *beta = nullptr;  // NFIELDS
    cplx<double> *hdev = nullptr;  // NFIELDS x (m+1)
    cplx<double> *ydev = nullptr;  // NFIELDS x m
    cplx<double> *cpart = nullptr; // NFIELDS x (m+1) x VBLOCKS
    double *rpart = nullptr;       // NFIELDS x VBLOCKS (x 8 for correlators)
    double *nrm = nullptr;         // NFIELDS
    int *active = nullptr;         // NFIELDS
    int *nused = nullptr;          // NFIELDS
    cplx<double> *hbuf = nullptr;  // pinned host: NFIELDS x (m+1)
    double *nbuf = nullptr;        // pinned host: 64
    void release()
    {
        void *p[] = {basis, b, x, t, res, beta, hdev, ydev, cpart, rpart, nrm, active, nused};
        for (void *q : p)
            if (q) (void)hipFree(q);
        if (hbuf) (void)hipHostFree(hbuf);
        if (nbuf) (void)hipHostFree(nbuf);
        basis = b = x = t = res = beta = hdev = ydev = cpart = nullptr;
        rpart = nrm = nullptr;
        active = nused = nullptr;
        hbuf = nullptr;
        nbuf = nullptr;
        vol = 0;
        m = 0;
    }
    int reserve(int64_t v, int mm)
    {
        if (v == vol && mm == m) return NRQ_OK;
        release();
        const size_t fb = (size_t)3 * (size_t)v * sizeof(cplx<double>);
        bool ok = hipMalloc(&basis, fb * NFIELDS * (size_t)(mm + 1)) == hipSuccess &&
                  hipMalloc(&b, fb * NFIELDS) == hipSuccess && hipMalloc(&x, fb * NFIELDS) == hipSuccess &&
                  hipMalloc(&t, fb * NFIELDS) == hipSuccess && hipMalloc(&res, fb * NFIELDS) == hipSuccess &&
                  hipMalloc(&beta, sizeof(cplx<double>) * NFIELDS) == hipSuccess &&
                  hipMalloc(&hdev, sizeof(cplx<double>) * NFIELDS * (mm + 1)) == hipSuccess &&
                  hipMalloc(&ydev, sizeof(cplx<double>) * NFIELDS * (mm + 1)) == hipSuccess &&
                  hipMalloc(&cpart, sizeof(cplx<double>) * NFIELDS * (mm + 1) * VBLOCKS) == hipSuccess &&
                  hipMalloc(&rpart, sizeof(double) * NFIELDS * VBLOCKS * 8) == hipSuccess &&
                  hipMalloc(&nrm, sizeof(double) * NFIELDS) == hipSuccess &&
                  hipMalloc(&active, sizeof(int) * NFIELDS) == hipSuccess &&
                  hipMalloc(&nused, sizeof(int) * NFIELDS) == hipSuccess &&
                  hipHostMalloc((void **)&hbuf, sizeof(cplx<double>) * NFIELDS * (mm + 1), hipHostMallocDefault) == hipSuccess &&
                  hipHostMalloc((void **)&nbuf, sizeof(double) * 64, hipHostMallocDefault) == hipSuccess;
        if (!ok) {
            release();
            return fail(NRQ_ERR_ALLOC, "CN: workspace allocation failed (V=%lld, m=%d)", (long long)v, mm);
        }
        vol = v;
        m = mm;
        return NRQ_OK;
    }
};
static cn_workspace g_ws;

static int upload_active(const std::vector<int> &act)
{
    NRQ_HIP(hipMemcpyAsync(g_ws.active, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice, stream()));
    return NRQ_OK;
}

// Applies y_r = x_r + beta_r K x_r for the active fields r < nf.
static int apply_op(const nrq_gauge *gf, const cplx<double> *in, cplx<double> *out, int nf)
{
    const geom g = lat().g;
    const dim3 grd((unsigned)((g.vol + 255) / 256), (nf + RCHUNK - 1) / RCHUNK);
    if (gf->precision == 64)
        hipLaunchKernelGGL(k_stencil<double>, grd, dim3(256), 0, stream(), (const cplx<double> *)gauge_links(gf), in, out,
                           g_ws.beta, g_ws.active, nf, g);
    else
        hipLaunchKernelGGL(k_stencil<float>, grd, dim3(256), 0, stream(), (const cplx<float> *)gauge_links(gf), in, out,
                           g_ws.beta, g_ws.active, nf, g);
    NRQ_HIP(hipGetLastError());
    return NRQ_OK;
}

// Squared norms of the active fields r < nf of v into host array out[nf].
static int field_norms2(const cplx<double> *v, double *out, int nf)
{
    const int64_t n = 3 * g_ws.vol;
    hipLaunchKernelGGL(k_norm2, dim3(VBLOCKS, nf), dim3(VTHREADS), 0, stream(), v, n, g_ws.active, g_ws.rpart);
    hipLaunchKernelGGL(k_fold_r, dim3(1), dim3(64), 0, stream(), g_ws.rpart, VBLOCKS, nf, g_ws.active, g_ws.nrm);
    NRQ_HIP(hipGetLastError());
    NRQ_HIP(hipMemcpyAsync(g_ws.nbuf, g_ws.nrm, sizeof(double) * nf, hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    memcpy(out, g_ws.nbuf, sizeof(double) * nf);
    return NRQ_OK;
}

// Batched restarted GMRES(m), classical Gram-Schmidt, zero initial guess,
// on the fields flagged in `act`: solves (1 + conj(alpha_r) K) x_r = b_r to
// ||b_r - A x_r|| <= tol ||b_r|| (FP64 true residual).
static int gmres_batched(const nrq_gauge *gf, int R, const double *bnorm2, nrq_solve_stats *st)
{
    const int m = g_ws.m;
    const int64_t n = 3 * g_ws.vol;
    const size_t fbytes = sizeof(cplx<double>) * (size_t)n;
    typedef std::complex<double> zc;
    std::vector<int> act(R, 1);
    std::vector<double> bn(R), rn(R);
    std::vector<int64_t> its(R, 0);
    for (int r = 0; r < R; ++r) bn[r] = std::sqrt(bnorm2[r]);
    NRQ_HIP(hipMemsetAsync(g_ws.x, 0, fbytes * R, stream()));
    // r0 = b
    NRQ_HIP(hipMemcpyAsync(g_ws.res, g_ws.b, fbytes * R, hipMemcpyDeviceToDevice, stream()));
    for (int r = 0; r < R; ++r) rn[r] = bn[r];
    std::vector<std::vector<zc>> H(R, std::vector<zc>((size_t)(m + 1) * m)), G(R, std::vector<zc>(m + 1)),
        SN(R, std::vector<zc>(m)), Y(R, std::vector<zc>(m));
    std::vector<std::vector<double>> CS(R, std::vector<double>(m));
    std::vector<int> jused(R, 0);
    int64_t total_its = 0, restarts = 0;
    double worst = 0;
    int status = NRQ_OK;
    std::vector<int> cycle(R);
    while (true) {
        bool any = false;
        for (int r = 0; r < R; ++r) any = any || act[r];
        if (!any) break;
        ++restarts;
        // v0 = res / ||res||
        for (int r = 0; r < R; ++r) {
            g_ws.nbuf[r] = act[r] ? rn[r] * rn[r] : 0.0;
            std::fill(G[r].begin(), G[r].end(), zc(0));
            G[r][0] = rn[r];
            jused[r] = 0;
        }
        NRQ_HIP(hipMemcpyAsync(g_ws.nrm, g_ws.nbuf, sizeof(double) * R, hipMemcpyHostToDevice, stream()));
        if (int e = upload_active(act)) return e;
        hipLaunchKernelGGL(k_scale, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), g_ws.res, g_ws.basis, n, g_ws.nrm, g_ws.active);
        cycle = act;
        for (int j = 0; j < m; ++j) {
            bool busy = false;
            for (int r = 0; r < R; ++r) busy = busy || cycle[r];
            if (!busy) break;
            if (int e = upload_active(cycle)) return e;
            cplx<double> *vj = g_ws.basis + (size_t)j * R * n;
            cplx<double> *w = g_ws.basis + (size_t)(j + 1) * R * n;
            if (int e = apply_op(gf, vj, w, R)) return e;
            const int nv = j + 1;
            hipLaunchKernelGGL(k_mdot, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), g_ws.basis, nv, w, n, R, g_ws.active, g_ws.cpart);
            hipLaunchKernelGGL(k_fold_c, dim3((nv * R + 63) / 64), dim3(64), 0, stream(), g_ws.cpart, VBLOCKS, nv, R, g_ws.active, g_ws.hdev);
            hipLaunchKernelGGL(k_cgs_update, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), g_ws.basis, nv, w, n, R, g_ws.active, g_ws.hdev, nv, g_ws.rpart);
            hipLaunchKernelGGL(k_fold_r, dim3(1), dim3(64), 0, stream(), g_ws.rpart, VBLOCKS, R, g_ws.active, g_ws.nrm);
            hipLaunchKernelGGL(k_scale, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), w, w, n, g_ws.nrm, g_ws.active);
            NRQ_HIP(hipGetLastError());
            NRQ_HIP(hipMemcpyAsync(g_ws.hbuf, g_ws.hdev, sizeof(cplx<double>) * nv * R, hipMemcpyDeviceToHost, stream()));
            NRQ_HIP(hipMemcpyAsync(g_ws.nbuf, g_ws.nrm, sizeof(double) * R, hipMemcpyDeviceToHost, stream()));
            NRQ_HIP(hipStreamSynchronize(stream()));
            ++total_its;
            for (int r = 0; r < R; ++r) {
                if (!cycle[r]) continue;
                ++its[r];
                zc *Hc = H[r].data() + (size_t)(m + 1) * j;
                for (int i = 0; i < nv; ++i) Hc[i] = zc(g_ws.hbuf[r * nv + i].re, g_ws.hbuf[r * nv + i].im);
                const double hn = std::sqrt(std::max(g_ws.nbuf[r], 0.0));
                for (int i = 0; i < j; ++i) {
                    const zc a = Hc[i], c = Hc[i + 1];
                    Hc[i] = CS[r][i] * a + SN[r][i] * c;
                    Hc[i + 1] = -std::conj(SN[r][i]) * a + CS[r][i] * c;
                }
                const zc a = Hc[j];
                const double den = std::sqrt(std::norm(a) + hn * hn);
                if (den == 0) {
                    CS[r][j] = 1;
                    SN[r][j] = 0;
                } else if (std::abs(a) == 0) {
                    CS[r][j] = 0;
                    SN[r][j] = 1;
                } else {
                    CS[r][j] = std::abs(a) / den;
                    SN[r][j] = (a / std::abs(a)) * (hn / den);
                }
                Hc[j] = CS[r][j] * a + SN[r][j] * hn;
                G[r][j + 1] = -std::conj(SN[r][j]) * G[r][j];
                G[r][j] = CS[r][j] * G[r][j];
                jused[r] = j + 1;
                if (std::abs(G[r][j + 1]) <= g_cn.tol * bn[r] || its[r] >= g_cn.maxit) cycle[r] = 0;
            }
        }
        // y = H^-1 g per field, x += V y
        std::vector<int> nu(R, 0);
        for (int r = 0; r < R; ++r) {
            if (!act[r]) continue;
            const int jj = jused[r];
            nu[r] = jj;
            for (int i = jj - 1; i >= 0; --i) {
                zc s = G[r][i];
                for (int k = i + 1; k < jj; ++k) s -= H[r][(size_t)(m + 1) * k + i] * Y[r][k];
                Y[r][i] = s / H[r][(size_t)(m + 1) * i + i];
            }
            for (int i = 0; i < jj; ++i) g_ws.hbuf[r * (m + 1) + i] = {Y[r][i].real(), Y[r][i].imag()};
        }
        if (int e = upload_active(act)) return e;
        NRQ_HIP(hipMemcpyAsync(g_ws.ydev, g_ws.hbuf, sizeof(cplx<double>) * R * (m + 1), hipMemcpyHostToDevice, stream()));
        NRQ_HIP(hipMemcpyAsync(g_ws.nused, nu.data(), sizeof(int) * R, hipMemcpyHostToDevice, stream()));
        hipLaunchKernelGGL(k_xupdate, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), g_ws.x, g_ws.basis, R, n, g_ws.ydev, m + 1, g_ws.nused, g_ws.active);
        // true residual
        if (int e = apply_op(gf, g_ws.x, g_ws.t, R)) return e;
        hipLaunchKernelGGL(k_residual, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), g_ws.b, g_ws.t, g_ws.res, n, g_ws.active, g_ws.rpart);
        hipLaunchKernelGGL(k_fold_r, dim3(1), dim3(64), 0, stream(), g_ws.rpart, VBLOCKS, R, g_ws.active, g_ws.nrm);
        NRQ_HIP(hipGetLastError());
        NRQ_HIP(hipMemcpyAsync(g_ws.nbuf, g_ws.nrm, sizeof(double) * R, hipMemcpyDeviceToHost, stream()));
        NRQ_HIP(hipStreamSynchronize(stream()));
        for (int r = 0; r < R; ++r) {
            if (!act[r]) continue;
            rn[r] = std::sqrt(std::max(g_ws.nbuf[r], 0.0));
            const double rel = rn[r] / bn[r];
            if (rn[r] <= g_cn.tol * bn[r]) {
                act[r] = 0;
                worst = std::max(worst, rel);
            } else if (its[r] >= g_cn.maxit) {
                act[r] = 0;
                worst = std::max(worst, rel);
                status = fail(NRQ_ERR_NOCONV, "CN: GMRES reached maxit=%lld with true relative residual %.3e",
                              (long long)g_cn.maxit, rel);
            }
        }
    }
    if (st) {
        int64_t s = 0;
        for (int r = 0; r < R; ++r) s += its[r];
        st->iterations += s;
        st->restarts += restarts;
        st->max_true_relres = std::max(st->max_true_relres, worst);
    }
    (void)total_its;
    return status;
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
    g_cn.tol = ksp_tol;
    for (int k = 0; k < 4; ++k) g_cn.c[k] = std::complex<double>(c[2 * k], c[2 * k + 1]);
    g_cn.initialised = true;
    return NRQ_OK;
}

extern "C" int nrq_cn_finalize(void)
{
    if (!g_cn.initialised) return fail(NRQ_ERR_STATE, "nrqcd: not initialised");
    if (stream()) (void)hipStreamSynchronize(stream());
    g_ws.release();
    g_cn.initialised = false;
    return NRQ_OK;
}

extern "C" int nrq_cn_set_solver(int kind, int restart, int64_t maxit)
{
    if (!g_cn.initialised) return fail(NRQ_ERR_STATE, "nrqcd: not initialised");
    if (kind != NRQ_SOLVER_GMRES_FP64) return fail(NRQ_ERR_UNSUPPORTED, "nrqcd: solver kind %d not available", kind);
    if (restart < 1 || restart > 200 || maxit < 1) return fail(NRQ_ERR_ARG, "nrqcd: bad restart or maxit");
    g_cn.kind = kind;
    g_cn.restart = restart;
    g_cn.maxit = maxit;
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
    for (int p = 0; p < 2; ++p) {
        const int D = dofs[p];
        const int col = D, c = D / 2, sp = D % 2;
        const int field = 12 * p + 2 * col + sp;
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
        for (int64_t s = 0; s < v; ++s)
            for (int j = 0; j < 6; ++j)
                for (int d = 0; d < 6; ++d) {
                    const int c = d / 2, sp = d % 2;
                    const double *z = src[p] + 2 * (d + 6 * j + 36 * s);
                    buf[((size_t)(2 * j + sp) * 3 + c) * (size_t)v + (size_t)s] = {z[0], z[1]};
                }
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

extern "C" int nrq_heavy_update(nrq_heavy *h, const nrq_gauge *gf, double mass, const double c[8], double stepwidth,
                                nrq_solve_stats *stats)
{
    if (!g_cn.initialised) return fail(NRQ_ERR_STATE, "nrqcd: not initialised");
    NEED_HEAVY(h);
    if (!gf || !gf->u[0]) return fail(NRQ_ERR_ARG, "heavy: null gauge configuration");
    if (gf->vol != h->vol) return fail(NRQ_ERR_ARG, "heavy: gauge and heavy field lattices differ");
    if (!c) return fail(NRQ_ERR_ARG, "heavy: null Wilson coefficients");
    if (!std::isfinite(mass) || mass == 0.0 || !std::isfinite(stepwidth)) return fail(NRQ_ERR_ARG, "heavy: bad mass or step width");
    for (int k = 1; k < 4; ++k)
        if (c[2 * k] != 0.0 || c[2 * k + 1] != 0.0)
            return fail(NRQ_ERR_UNSUPPORTED, "heavy: only the kinetic term c1 is implemented (c%d != 0)", k + 1);
    if (int r = g_ws.reserve(h->vol, g_cn.restart)) return r;
    nrq_solve_stats st = {0, 0, 0, 0.0};
    const double a0 = lat().g.a[0];
    const std::complex<double> c1(c[0], c[1]);
    const std::complex<double> mi(0.0, -(stepwidth / 2.0 * a0));
    const std::complex<double> alpha_q = mi * c1 / mass;
    const std::complex<double> alpha_a = mi * std::conj(c1) / (-mass);
    std::vector<cplx<double>> beta(NFIELDS);
    std::vector<int> all(NFIELDS, 1);
    // rhs = S psi (SystMat_herm before the in-place Hermitian transpose) for
    // all 24 fields, into the scratch slot array t
    for (int r = 0; r < NFIELDS; ++r) {
        const std::complex<double> a = r < 12 ? alpha_q : alpha_a;
        beta[r] = {a.real(), a.imag()};
    }
    NRQ_HIP(hipMemcpyAsync(g_ws.beta, beta.data(), sizeof(cplx<double>) * NFIELDS, hipMemcpyHostToDevice, stream()));
    if (int e = upload_active(all)) return e;
    if (int e = apply_op(gf, (const cplx<double> *)h->psi, g_ws.t, NFIELDS)) return e;
    double bn2_all[NFIELDS];
    if (int e = field_norms2(g_ws.t, bn2_all, NFIELDS)) return e;
    // compact the non-zero right-hand sides into slots 0..R-1 (a zero
    // right-hand side has the zero solution, as KSPSolve returns for b = 0)
    std::vector<int> field;
    for (int r = 0; r < NFIELDS; ++r)
        if (bn2_all[r] > 0.0) field.push_back(r);
    const int R = (int)field.size();
    st.solves = R;
    const size_t fb = (size_t)3 * (size_t)h->vol * sizeof(cplx<double>);
    double bn2[NFIELDS];
    std::vector<cplx<double>> sbeta(R > 0 ? R : 1);
    for (int q = 0; q < R; ++q) {
        const int r = field[q];
        NRQ_HIP(hipMemcpyAsync((char *)g_ws.b + fb * q, (char *)g_ws.t + fb * r, fb, hipMemcpyDeviceToDevice, stream()));
        bn2[q] = bn2_all[r];
        // operator S^+ = 1 + conj(alpha) K
        const std::complex<double> a = std::conj(r < 12 ? alpha_q : alpha_a);
        sbeta[q] = {a.real(), a.imag()};
    }
    int status = NRQ_OK;
    if (R > 0) {
        NRQ_HIP(hipMemcpyAsync(g_ws.beta, sbeta.data(), sizeof(cplx<double>) * R, hipMemcpyHostToDevice, stream()));
        status = gmres_batched(gf, R, bn2, &st);
    }
    // scatter the solutions back; fields with a zero right-hand side are zero
    for (int r = 0; r < NFIELDS; ++r)
        if (!(bn2_all[r] > 0.0)) NRQ_HIP(hipMemsetAsync((char *)h->psi + fb * r, 0, fb, stream()));
    for (int q = 0; q < R; ++q)
        NRQ_HIP(hipMemcpyAsync((char *)h->psi + fb * field[q], (char *)g_ws.x + fb * q, fb, hipMemcpyDeviceToDevice, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    if (stats) *stats = st;
    return status;
}

extern "C" int nrq_heavy_norms(const nrq_heavy *h, double *nq, double *na)
{
    NEED_HEAVY(h);
    if (!nq || !na) return fail(NRQ_ERR_ARG, "null argument");
    if (int r = g_ws.reserve(h->vol, g_cn.initialised ? g_cn.restart : 30)) return r;
    std::vector<int> all(NFIELDS, 1);
    if (int e = upload_active(all)) return e;
    double n2[NFIELDS];
    if (int e = field_norms2((const cplx<double> *)h->psi, n2, NFIELDS)) return e;
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
    if (int r = g_ws.reserve(h->vol, g_cn.initialised ? g_cn.restart : 30)) return r;
    const int nb = 256;
    hipLaunchKernelGGL(k_corr, dim3(nb), dim3(REDUCE_THREADS), 0, stream(), (const cplx<double> *)h->psi, h->vol, g_ws.rpart);
    std::vector<int> all(NFIELDS, 1);
    if (int e = upload_active(all)) return e;
    hipLaunchKernelGGL(k_fold_r, dim3(1), dim3(64), 0, stream(), g_ws.rpart, nb, 8, g_ws.active, g_ws.nrm);
    NRQ_HIP(hipGetLastError());
    NRQ_HIP(hipMemcpyAsync(g_ws.nbuf, g_ws.nrm, sizeof(double) * 8, hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    memcpy(out, g_ws.nbuf, sizeof(double) * 8);
    return NRQ_OK;
}
