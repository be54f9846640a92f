// nrq_krylov.hip -- batched Krylov solver for the NRQCD Crank-Nicolson systems
// (replaces PETSc KSP, nrqcd.f90:786-804, 1103-1105, 1123, 1161).
//
// Two drivers share one GMRES(m) core with classical Gram-Schmidt:
//  * KRYLOV_GMRES_FP64: GMRES(m) in FP64 from a zero guess (PETSc's defaults,
//    judged on the FP64 true residual).
//  * KRYLOV_GMRES_MIXED: iterative refinement.  The FP64 true residual
//    r = y - A x is formed with the FP64 operator; the correction A d = r is
//    solved by GMRES in FP32 (FP32 copy of the links, FP32 vectors) to
//    inner_tol; x += d; repeat until ||r|| <= tol ||y||.  The FP64 residual
//    is the judge, so FP32 changes the iteration count, never the answer.
//
// All slots advance together: one stencil launch per iteration streams the
// links once per chunk of RCHUNK slots; the host keeps the small Hessenberg
// systems (one Givens-rotated least-squares problem per slot).
#include "nrq_krylov.h"
#include "nrq_map.h"
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>

namespace nrq {

const void *gauge_links(const nrq_gauge *h);

constexpr int RCHUNK = 4;     // slots per thread in the stencil (link reuse)
constexpr int CHEB_MAX = 64;  // Chebyshev steps per coefficient table
constexpr int VBLOCKS = 256;  // blocks per slot in the vector kernels
constexpr int VTHREADS = 256;

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// out_r = in_r + beta_r K in_r for the active slots of one chunk; links of
// type RL are converted to the compute type T on load and reused across the
// RCHUNK slots of the chunk.
// RESID: out_r = b_r - (in_r + beta_r K in_r) instead.  map (may be null)
// gives the input field of slot r (in + map[r] * 3V).
template <typename RL, typename T, bool RESID>
__global__ __launch_bounds__(256) void k_stencil(const cplx<RL> *__restrict__ u, const cplx<T> *__restrict__ in,
                                                 cplx<T> *__restrict__ out, const cplx<double> *__restrict__ beta,
                                                 const int *__restrict__ active, int nf, geom g, lf_map mp,
                                                 const int *__restrict__ map, const cplx<T> *__restrict__ b)
{
    int64_t s;
    int chunk;
    if (!lf_site(mp, g, s, chunk)) return;
    const int r0 = chunk * RCHUNK;
    const int64_t fs = 3 * g.vol;
    cplx<T> acc[RCHUNK][3];
    bool on[RCHUNK];
#pragma unroll
    for (int q = 0; q < RCHUNK; ++q) {
        on[q] = (r0 + q < nf) && active[r0 + q];
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[q][c] = {T(0), T(0)};
    }
    for (int i = 0; i < 3; ++i) {
        const T ia2 = T(1.0 / (g.a[i + 1] * g.a[i + 1]));
        const int64_t sp = neib_up(g, s, i), sm = neib_dn(g, s, i);
        cplx<T> up[9], um[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            const cplx<RL> a = u[((int64_t)i * 9 + e) * g.vol + s];
            const cplx<RL> b = u[((int64_t)i * 9 + e) * g.vol + sm];
            up[e] = {(T)a.re, (T)a.im};
            um[e] = {(T)b.re, (T)b.im};
        }
#pragma unroll
        for (int q = 0; q < RCHUNK; ++q) {
            if (!on[q]) continue;
            const cplx<T> *f = in + (int64_t)(map ? map[r0 + q] : r0 + q) * fs;
            cplx<T> vp[3], vm[3], v0[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                vp[c] = f[(int64_t)c * g.vol + sp];
                vm[c] = f[(int64_t)c * g.vol + sm];
                v0[c] = f[(int64_t)c * g.vol + s];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                // U_i(x) psi(x+i) + U_i(x-i)^+ psi(x-i)
                cplx<T> hop = up[3 * c] * vp[0] + up[3 * c + 1] * vp[1] + up[3 * c + 2] * vp[2];
                hop = hop + cmul(um[c], vm[0]) + cmul(um[3 + c], vm[1]) + cmul(um[6 + c], vm[2]);
                acc[q][c] = acc[q][c] + ia2 * (v0[c] - T(0.5) * hop);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RCHUNK; ++q) {
        if (!on[q]) continue;
        const int r = r0 + q;
        const cplx<T> bt = {(T)beta[r].re, (T)beta[r].im};
        const cplx<T> *f = in + (int64_t)(map ? map[r] : r) * fs;
        cplx<T> *o = out + (int64_t)r * fs;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const cplx<T> ax = f[(int64_t)c * g.vol + s] + bt * acc[q][c];
            o[(int64_t)c * g.vol + s] = RESID ? b[(int64_t)r * fs + (int64_t)c * g.vol + s] - ax : ax;
        }
    }
}

// One Chebyshev step on the segment [1, 1 + beta_r kmax] of the spectrum of
// A_r = 1 + beta_r K, fused with the operator application:
//   x_{k+1} = rho_r (x_k + (b_r - A_r x_k) / c_r) + (1 - rho_r) x_{k-1}
// c_r = 1 + beta_r kmax / 2 is the segment centre; x_{k-1} is read and
// x_{k+1} written at the same site of `xio` (first step: rho = 1, no read).
template <typename RL, typename T>
__global__ __launch_bounds__(256) void k_cheb(const cplx<RL> *__restrict__ u, const cplx<T> *__restrict__ xk,
                                              cplx<T> *__restrict__ xio, const cplx<T> *__restrict__ b,
                                              const cplx<double> *__restrict__ beta, const cplx<double> *__restrict__ coef,
                                              const int *__restrict__ active, int nf, geom g, lf_map mp, int first)
{
    int64_t s;
    int chunk;
    if (!lf_site(mp, g, s, chunk)) return;
    const int r0 = chunk * RCHUNK;
    const int64_t fs = 3 * g.vol;
    cplx<T> acc[RCHUNK][3];
    bool on[RCHUNK];
#pragma unroll
    for (int q = 0; q < RCHUNK; ++q) {
        on[q] = (r0 + q < nf) && active[r0 + q];
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[q][c] = {T(0), T(0)};
    }
    for (int i = 0; i < 3; ++i) {
        const T ia2 = T(1.0 / (g.a[i + 1] * g.a[i + 1]));
        const int64_t sp = neib_up(g, s, i), sm = neib_dn(g, s, i);
        cplx<T> up[9], um[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            const cplx<RL> a = u[((int64_t)i * 9 + e) * g.vol + s];
            const cplx<RL> bb = u[((int64_t)i * 9 + e) * g.vol + sm];
            up[e] = {(T)a.re, (T)a.im};
            um[e] = {(T)bb.re, (T)bb.im};
        }
#pragma unroll
        for (int q = 0; q < RCHUNK; ++q) {
            if (!on[q]) continue;
            const cplx<T> *f = xk + (int64_t)(r0 + q) * fs;
            cplx<T> vp[3], vm[3], v0[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                vp[c] = f[(int64_t)c * g.vol + sp];
                vm[c] = f[(int64_t)c * g.vol + sm];
                v0[c] = f[(int64_t)c * g.vol + s];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                cplx<T> hop = up[3 * c] * vp[0] + up[3 * c + 1] * vp[1] + up[3 * c + 2] * vp[2];
                hop = hop + cmul(um[c], vm[0]) + cmul(um[3 + c], vm[1]) + cmul(um[6 + c], vm[2]);
                acc[q][c] = acc[q][c] + ia2 * (v0[c] - T(0.5) * hop);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RCHUNK; ++q) {
        if (!on[q]) continue;
        const int r = r0 + q;
        const cplx<T> bt = {(T)beta[r].re, (T)beta[r].im};
        const cplx<T> ic = {(T)coef[2 * r].re, (T)coef[2 * r].im};    // 1 / c_r
        const cplx<T> rho = {(T)coef[2 * r + 1].re, (T)coef[2 * r + 1].im};
        const cplx<T> one_m_rho = {T(1) - rho.re, -rho.im};
        const int64_t base = (int64_t)r * fs + s;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int64_t at = base + (int64_t)c * g.vol;
            const cplx<T> x0 = xk[at];
            const cplx<T> res = b[at] - (x0 + bt * acc[q][c]);
            const cplx<T> nx = rho * (x0 + ic * res);
            xio[at] = first ? nx : nx + one_m_rho * xio[at];
        }
    }
}

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

// partial[r][i][blk] = this block's share of <V_i,r, w_r>, i < nv (FP64 sums)
template <typename T>
__global__ __launch_bounds__(VTHREADS) void k_mdot(const cplx<T> *__restrict__ basis, int nv, const cplx<T> *__restrict__ w,
                                                   int64_t n, int nf, const int *__restrict__ active,
                                                   cplx<double> *__restrict__ partial)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    __shared__ double sh[8];
    const cplx<T> *wr = w + (int64_t)r * n;
    for (int i = 0; i < nv; ++i) {
        const cplx<T> *vr = basis + ((int64_t)i * nf + r) * n;
        double re = 0, im = 0;
        for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
            const cplx<T> a = vr[k], b = wr[k];
            re += (double)a.re * b.re + (double)a.im * b.im;
            im += (double)a.re * b.im - (double)a.im * b.re;
        }
        re = block_sum(re, sh);
        im = block_sum(im, sh);
        if (threadIdx.x == 0) partial[((int64_t)r * nv + i) * gridDim.x + blockIdx.x] = {re, im};
    }
}

// out[r*count + i] = sum over blocks (fixed order)
__global__ void k_fold_c(const cplx<double> *__restrict__ partial, int nblk, int count, int nf,
                         const int *__restrict__ active, cplx<double> *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * nf) return;
    if (!active[t / count]) return;
    cplx<double> s = {0.0, 0.0};
    for (int b = 0; b < nblk; ++b) s = s + partial[(int64_t)t * nblk + b];
    out[t] = s;
}

// CGS update w_r -= sum_i h[r][i] V_i,r and partial ||w_r||^2
template <typename T>
__global__ __launch_bounds__(VTHREADS) void k_cgs_update(const cplx<T> *__restrict__ basis, int nv, cplx<T> *__restrict__ w,
                                                         int64_t n, int nf, const int *__restrict__ active,
                                                         const cplx<double> *__restrict__ h, double *__restrict__ partial)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    __shared__ double sh[8];
    cplx<T> *wr = w + (int64_t)r * n;
    double nrm = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const cplx<T> w0 = wr[k];
        cplx<double> a = {(double)w0.re, (double)w0.im};
        for (int i = 0; i < nv; ++i) {
            const cplx<T> v = basis[((int64_t)i * nf + r) * n + k];
            a = a - h[(int64_t)r * nv + i] * cplx<double>{(double)v.re, (double)v.im};
        }
        const cplx<T> st = {(T)a.re, (T)a.im};
        wr[k] = st;
        nrm += (double)st.re * st.re + (double)st.im * st.im;
    }
    nrm = block_sum(nrm, sh);
    if (threadIdx.x == 0) partial[(int64_t)r * gridDim.x + blockIdx.x] = nrm;
}

__global__ void k_fold_r(const double *__restrict__ partial, int nblk, int nf, const int *__restrict__ active,
                         double *__restrict__ out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nf || !active[r]) return;
    double s = 0;
    for (int b = 0; b < nblk; ++b) s += partial[(int64_t)r * nblk + b];
    out[r] = s;
}

// dst_r = src_r / sqrt(nrm2[r]) (zero norm leaves zeros)
template <typename T>
__global__ void k_scale(const cplx<T> *__restrict__ src, cplx<T> *__restrict__ dst, int64_t n,
                        const double *__restrict__ nrm2, const int *__restrict__ active)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    const double nn = nrm2[r];
    const T inv = (T)(nn > 0 ? 1.0 / sqrt(nn) : 0.0);
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        dst[(int64_t)r * n + k] = inv * src[(int64_t)r * n + k];
}

// x_r += sum_i y[r][i] V_i,r
template <typename T>
__global__ void k_xupdate(cplx<T> *__restrict__ x, const cplx<T> *__restrict__ basis, int nf, int64_t n,
                          const cplx<double> *__restrict__ y, int ystride, const int *__restrict__ nused,
                          const int *__restrict__ active)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    const int nv = nused[r];
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const cplx<T> x0 = x[(int64_t)r * n + k];
        cplx<double> a = {(double)x0.re, (double)x0.im};
        for (int i = 0; i < nv; ++i) {
            const cplx<T> v = basis[((int64_t)i * nf + r) * n + k];
            a = a + y[(int64_t)r * ystride + i] * cplx<double>{(double)v.re, (double)v.im};
        }
        x[(int64_t)r * n + k] = {(T)a.re, (T)a.im};
    }
}

// res_r = b_r - t_r and partial ||res_r||^2
template <typename T>
__global__ __launch_bounds__(VTHREADS) void k_residual(const cplx<T> *__restrict__ b, const cplx<T> *__restrict__ t,
                                                       cplx<T> *__restrict__ res, int64_t n, const int *__restrict__ active,
                                                       double *__restrict__ partial)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    __shared__ double sh[8];
    double nrm = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const cplx<T> a = b[(int64_t)r * n + k] - t[(int64_t)r * n + k];
        res[(int64_t)r * n + k] = a;
        nrm += (double)a.re * a.re + (double)a.im * a.im;
    }
    nrm = block_sum(nrm, sh);
    if (threadIdx.x == 0) partial[(int64_t)r * gridDim.x + blockIdx.x] = nrm;
}

template <typename T>
__global__ __launch_bounds__(VTHREADS) void k_norm2(const cplx<T> *__restrict__ x, int64_t n, const int *__restrict__ active,
                                                    double *__restrict__ partial)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    __shared__ double sh[8];
    double nrm = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const cplx<T> a = x[(int64_t)r * n + k];
        nrm += (double)a.re * a.re + (double)a.im * a.im;
    }
    nrm = block_sum(nrm, sh);
    if (threadIdx.x == 0) partial[(int64_t)r * gridDim.x + blockIdx.x] = nrm;
}

// dst32_r = src64_r / sqrt(nrm2[r])  (normalised FP32 right-hand side)
__global__ void k_demote(const cplx<double> *__restrict__ src, cplx<float> *__restrict__ dst, int64_t n,
                         const double *__restrict__ nrm2, const int *__restrict__ active)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    const double inv = nrm2[r] > 0 ? 1.0 / sqrt(nrm2[r]) : 0.0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const cplx<double> a = src[(int64_t)r * n + k];
        dst[(int64_t)r * n + k] = {(float)(a.re * inv), (float)(a.im * inv)};
    }
}

// x64_r += sqrt(nrm2[r]) d32_r
__global__ void k_promote_add(cplx<double> *__restrict__ x, const cplx<float> *__restrict__ d, int64_t n,
                              const double *__restrict__ nrm2, const int *__restrict__ active)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    const double sc = sqrt(nrm2[r]);
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const cplx<float> a = d[(int64_t)r * n + k];
        x[(int64_t)r * n + k] = x[(int64_t)r * n + k] + sc * cplx<double>{(double)a.re, (double)a.im};
    }
}

__global__ void k_links_to_f32(const cplx<double> *__restrict__ u, cplx<float> *__restrict__ v, int64_t n)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        v[k] = {(float)u[k].re, (float)u[k].im};
}

// x_r = prev_r + c_r (y_r - prev_r): the first-order guess
// prev + (alpha - beta) K prev with K prev = (y - prev)/alpha
__global__ void k_guess(cplx<double> *__restrict__ x, const cplx<double> *__restrict__ prev, const cplx<double> *__restrict__ y,
                        const cplx<double> *__restrict__ coef, int64_t n, const int *__restrict__ active)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    const cplx<double> c = coef[r];
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t at = (int64_t)r * n + k;
        x[at] = prev[at] + c * (y[at] - prev[at]);
    }
}

// ---------------------------------------------------------------------------
// workspace
// ---------------------------------------------------------------------------
struct workspace {
    int64_t vol = 0;
    int m = 0;
    void *basis = nullptr;             // (m+1) x NSLOTS x 3V complex double (or float)
    cplx<double> *y = nullptr, *x = nullptr, *t = nullptr, *res = nullptr, *scr = nullptr;
    cplx<float> *y32 = nullptr, *x32 = nullptr, *t32 = nullptr, *res32 = nullptr;
    cplx<float> *links32 = nullptr;    // FP32 copy of the links for the inner solves
    cplx<double> *beta = nullptr, *coef = nullptr, *hdev = nullptr, *ydev = nullptr, *cpart = nullptr;
    cplx<double> *cheb = nullptr;      // Chebyshev coefficients [iteration][slot][2]
    double *rpart = nullptr, *nrm = nullptr;
    int *active = nullptr, *nused = nullptr;
    cplx<double> *hbuf = nullptr;      // pinned host
    double *nbuf = nullptr;            // pinned host (64 doubles)

    void release()
    {
        void *dev[] = {basis, y, x, t, res, scr, y32, x32, t32, res32, links32, beta, coef, hdev, ydev, cpart, rpart, nrm, active, nused, cheb};
        for (void *q : dev)
            if (q) (void)hipFree(q);
        if (hbuf) (void)hipHostFree(hbuf);
        if (nbuf) (void)hipHostFree(nbuf);
        *this = workspace();
    }

    int reserve(int64_t v, int mm)
    {
        if (v == vol && mm == m) return NRQ_OK;
        release();
        const size_t f64 = (size_t)3 * (size_t)v * sizeof(cplx<double>) * NSLOTS;
        const size_t f32 = f64 / 2;
        auto dm = [](void **p, size_t b) { return hipMalloc(p, b) == hipSuccess; };
        bool ok = dm(&basis, f64 * (size_t)(mm + 1)) && dm((void **)&y, f64) && dm((void **)&x, f64) &&
                  dm((void **)&t, f64) && dm((void **)&res, f64) && dm((void **)&scr, f64) && dm((void **)&y32, f32) &&
                  dm((void **)&x32, f32) && dm((void **)&t32, f32) && dm((void **)&res32, f32) &&
                  dm((void **)&links32, (size_t)27 * (size_t)v * sizeof(cplx<float>)) &&
                  dm((void **)&beta, sizeof(cplx<double>) * NSLOTS) && dm((void **)&coef, sizeof(cplx<double>) * NSLOTS) &&
                  dm((void **)&cheb, sizeof(cplx<double>) * 2 * NSLOTS * CHEB_MAX) &&
                  dm((void **)&hdev, sizeof(cplx<double>) * NSLOTS * (mm + 1)) &&
                  dm((void **)&ydev, sizeof(cplx<double>) * NSLOTS * (mm + 1)) &&
                  dm((void **)&cpart, sizeof(cplx<double>) * NSLOTS * (mm + 1) * VBLOCKS) &&
                  dm((void **)&rpart, sizeof(double) * NSLOTS * VBLOCKS * 8) && dm((void **)&nrm, sizeof(double) * 64) &&
                  dm((void **)&active, sizeof(int) * NSLOTS) && dm((void **)&nused, sizeof(int) * NSLOTS) &&
                  hipHostMalloc((void **)&hbuf, sizeof(cplx<double>) * NSLOTS * (mm + 1), hipHostMallocDefault) == hipSuccess &&
                  hipHostMalloc((void **)&nbuf, sizeof(double) * 64, hipHostMallocDefault) == hipSuccess;
        if (!ok) {
            (void)hipGetLastError();
            release();
            return fail(NRQ_ERR_ALLOC, "CN: workspace allocation failed (V=%lld, m=%d)", (long long)v, mm);
        }
        vol = v;
        m = mm;
        return NRQ_OK;
    }
};
static workspace g_ws;

int krylov_reserve(int64_t vol, int m) { return g_ws.reserve(vol, m); }
void krylov_release() { g_ws.release(); }
cplx<double> *krylov_rhs() { return g_ws.y; }
cplx<double> *krylov_solution() { return g_ws.x; }
cplx<double> *krylov_scratch() { return g_ws.scr; }
double *krylov_reduce_scratch() { return g_ws.rpart; }
double *krylov_host_scratch() { return g_ws.nbuf; }

static int upload_active(const std::vector<int> &act)
{
    NRQ_HIP(hipMemcpyAsync(g_ws.active, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice, stream()));
    return NRQ_OK;
}

static int upload_beta(const cplx<double> *beta, int nf)
{
    NRQ_HIP(hipMemcpyAsync(g_ws.beta, beta, sizeof(cplx<double>) * nf, hipMemcpyHostToDevice, stream()));
    return NRQ_OK;
}

template <typename RL, typename T>
static int apply_t(const cplx<RL> *u, const cplx<T> *in, cplx<T> *out, int nf, const int *map = nullptr,
                   const cplx<T> *resid_b = nullptr)
{
    const geom g = lat().g;
    const lf_map mp = make_lf_map(g, 256, 8, 1, (nf + RCHUNK - 1) / RCHUNK);
    if (resid_b)
        hipLaunchKernelGGL((k_stencil<RL, T, true>), dim3((unsigned)(mp.ngroup * mp.nsb)), dim3(256), 0, stream(), u, in,
                           out, g_ws.beta, g_ws.active, nf, g, mp, map, resid_b);
    else
        hipLaunchKernelGGL((k_stencil<RL, T, false>), dim3((unsigned)(mp.ngroup * mp.nsb)), dim3(256), 0, stream(), u, in,
                           out, g_ws.beta, g_ws.active, nf, g, mp, map, (const cplx<T> *)nullptr);
    NRQ_HIP(hipGetLastError());
    return NRQ_OK;
}

// FP64 operator with the configuration's own links (FP64 or FP32 storage)
static int apply64(const nrq_gauge *gf, const cplx<double> *in, cplx<double> *out, int nf)
{
    if (gf->precision == 64) return apply_t<double, double>((const cplx<double> *)gauge_links(gf), in, out, nf);
    return apply_t<float, double>((const cplx<float> *)gauge_links(gf), in, out, nf);
}

template <typename T> static int norms2_t(const cplx<T> *v, int nf, double *out)
{
    const int64_t n = 3 * g_ws.vol;
    hipLaunchKernelGGL(k_norm2<T>, dim3(VBLOCKS, nf), dim3(VTHREADS), 0, stream(), v, n, g_ws.active, g_ws.rpart);
    hipLaunchKernelGGL(k_fold_r, dim3(1), dim3(64), 0, stream(), g_ws.rpart, VBLOCKS, nf, g_ws.active, g_ws.nrm);
    NRQ_HIP(hipGetLastError());
    NRQ_HIP(hipMemcpyAsync(g_ws.nbuf, g_ws.nrm, sizeof(double) * nf, hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    memcpy(out, g_ws.nbuf, sizeof(double) * nf);
    return NRQ_OK;
}

int krylov_apply(const nrq_gauge *gf, const cplx<double> *x, cplx<double> *y, int nf, const cplx<double> *beta)
{
    std::vector<int> all(nf, 1);
    if (int e = upload_active(all)) return e;
    if (int e = upload_beta(beta, nf)) return e;
    return apply64(gf, x, y, nf);
}

int krylov_norms2(const cplx<double> *x, int nf, double *out)
{
    std::vector<int> all(nf, 1);
    if (int e = upload_active(all)) return e;
    return norms2_t<double>(x, nf, out);
}

// ---------------------------------------------------------------------------
// GMRES(m) core: solves (1 + beta_r K) x_r = b_r for the slots with act[r],
// from the current x (zero when zero_guess), until the T-precision true
// residual satisfies ||b - A x|| <= tol ||b|| or its[r] reaches maxit.
// rel_out[r] receives the last true relative residual.  beta must already be
// on the device.
// ---------------------------------------------------------------------------
template <typename T, typename RL>
static int gmres_core(const cplx<RL> *links, int R, const cplx<T> *b, cplx<T> *x, cplx<T> *basis, cplx<T> *t,
                      cplx<T> *res, bool zero_guess, const double *bnorm, std::vector<int> act, double tol, int m,
                      int64_t maxit, std::vector<int64_t> &its, double *rel_out, int64_t &cycles)
{
    typedef std::complex<double> zc;
    const int64_t n = 3 * g_ws.vol;
    const size_t fbytes = sizeof(cplx<T>) * (size_t)n;
    std::vector<double> rn(R, 0.0);
    if (int e = upload_active(act)) return e;
    if (zero_guess) {
        NRQ_HIP(hipMemsetAsync(x, 0, fbytes * R, stream()));
        NRQ_HIP(hipMemcpyAsync(res, b, fbytes * R, hipMemcpyDeviceToDevice, stream()));
        for (int r = 0; r < R; ++r) rn[r] = bnorm[r];
    } else {
        if (int e = apply_t<RL, T>(links, x, t, R)) return e;
        hipLaunchKernelGGL(k_residual<T>, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), b, t, res, n, g_ws.active, g_ws.rpart);
        hipLaunchKernelGGL(k_fold_r, dim3(1), dim3(64), 0, stream(), g_ws.rpart, VBLOCKS, R, g_ws.active, g_ws.nrm);
        NRQ_HIP(hipMemcpyAsync(g_ws.nbuf, g_ws.nrm, sizeof(double) * R, hipMemcpyDeviceToHost, stream()));
        NRQ_HIP(hipStreamSynchronize(stream()));
        for (int r = 0; r < R; ++r) rn[r] = act[r] ? std::sqrt(std::max(g_ws.nbuf[r], 0.0)) : 0.0;
    }
    for (int r = 0; r < R; ++r) {
        if (!act[r]) continue;
        rel_out[r] = bnorm[r] > 0 ? rn[r] / bnorm[r] : 0.0;
        if (rn[r] <= tol * bnorm[r]) act[r] = 0;
    }
    std::vector<std::vector<zc>> H(R, std::vector<zc>((size_t)(m + 1) * m)), G(R, std::vector<zc>(m + 1)),
        SN(R, std::vector<zc>(m)), Y(R, std::vector<zc>(m));
    std::vector<std::vector<double>> CS(R, std::vector<double>(m));
    std::vector<int> jused(R, 0), cycle(R);
    while (true) {
        bool any = false;
        for (int r = 0; r < R; ++r) any = any || act[r];
        if (!any) break;
        ++cycles;
        for (int r = 0; r < R; ++r) {
            g_ws.nbuf[r] = act[r] ? rn[r] * rn[r] : 0.0;
            std::fill(G[r].begin(), G[r].end(), zc(0));
            G[r][0] = rn[r];
            jused[r] = 0;
        }
        NRQ_HIP(hipMemcpyAsync(g_ws.nrm, g_ws.nbuf, sizeof(double) * R, hipMemcpyHostToDevice, stream()));
        if (int e = upload_active(act)) return e;
        hipLaunchKernelGGL(k_scale<T>, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), (const cplx<T> *)res, basis, n, g_ws.nrm,
                           g_ws.active);
        cycle = act;
        for (int j = 0; j < m; ++j) {
            bool busy = false;
            for (int r = 0; r < R; ++r) busy = busy || cycle[r];
            if (!busy) break;
            if (int e = upload_active(cycle)) return e;
            cplx<T> *vj = basis + (size_t)j * R * n;
            cplx<T> *w = basis + (size_t)(j + 1) * R * n;
            if (int e = apply_t<RL, T>(links, vj, w, R)) return e;
            const int nv = j + 1;
            hipLaunchKernelGGL(k_mdot<T>, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), (const cplx<T> *)basis, nv,
                               (const cplx<T> *)w, n, R, g_ws.active, g_ws.cpart);
            hipLaunchKernelGGL(k_fold_c, dim3((nv * R + 63) / 64), dim3(64), 0, stream(), g_ws.cpart, VBLOCKS, nv, R, g_ws.active,
                               g_ws.hdev);
            hipLaunchKernelGGL(k_cgs_update<T>, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), (const cplx<T> *)basis, nv, w, n, R,
                               g_ws.active, g_ws.hdev, g_ws.rpart);
            hipLaunchKernelGGL(k_fold_r, dim3(1), dim3(64), 0, stream(), g_ws.rpart, VBLOCKS, R, g_ws.active, g_ws.nrm);
            hipLaunchKernelGGL(k_scale<T>, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), (const cplx<T> *)w, w, n, g_ws.nrm,
                               g_ws.active);
            NRQ_HIP(hipGetLastError());
            NRQ_HIP(hipMemcpyAsync(g_ws.hbuf, g_ws.hdev, sizeof(cplx<double>) * nv * R, hipMemcpyDeviceToHost, stream()));
            NRQ_HIP(hipMemcpyAsync(g_ws.nbuf, g_ws.nrm, sizeof(double) * R, hipMemcpyDeviceToHost, stream()));
            NRQ_HIP(hipStreamSynchronize(stream()));
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
                if (std::abs(G[r][j + 1]) <= tol * bnorm[r] || its[r] >= maxit) cycle[r] = 0;
            }
        }
        // y = H^-1 g per slot, x += V y
        std::vector<int> nu(R, 0);
        for (int r = 0; r < R; ++r) {
            if (!act[r]) continue;
            const int jj = jused[r];
            nu[r] = jj;
            for (int i = jj - 1; i >= 0; --i) {
                zc sacc = G[r][i];
                for (int k = i + 1; k < jj; ++k) sacc -= H[r][(size_t)(m + 1) * k + i] * Y[r][k];
                Y[r][i] = sacc / H[r][(size_t)(m + 1) * i + i];
            }
            for (int i = 0; i < jj; ++i) g_ws.hbuf[r * (m + 1) + i] = {Y[r][i].real(), Y[r][i].imag()};
        }
        if (int e = upload_active(act)) return e;
        NRQ_HIP(hipMemcpyAsync(g_ws.ydev, g_ws.hbuf, sizeof(cplx<double>) * R * (m + 1), hipMemcpyHostToDevice, stream()));
        NRQ_HIP(hipMemcpyAsync(g_ws.nused, nu.data(), sizeof(int) * R, hipMemcpyHostToDevice, stream()));
        hipLaunchKernelGGL(k_xupdate<T>, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), x, (const cplx<T> *)basis, R, n,
                           g_ws.ydev, m + 1, g_ws.nused, g_ws.active);
        // true residual in precision T
        if (int e = apply_t<RL, T>(links, x, t, R)) return e;
        hipLaunchKernelGGL(k_residual<T>, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), b, (const cplx<T> *)t, res, n,
                           g_ws.active, g_ws.rpart);
        hipLaunchKernelGGL(k_fold_r, dim3(1), dim3(64), 0, stream(), g_ws.rpart, VBLOCKS, R, g_ws.active, g_ws.nrm);
        NRQ_HIP(hipGetLastError());
        NRQ_HIP(hipMemcpyAsync(g_ws.nbuf, g_ws.nrm, sizeof(double) * R, hipMemcpyDeviceToHost, stream()));
        NRQ_HIP(hipStreamSynchronize(stream()));
        for (int r = 0; r < R; ++r) {
            if (!act[r]) continue;
            rn[r] = std::sqrt(std::max(g_ws.nbuf[r], 0.0));
            rel_out[r] = rn[r] / bnorm[r];
            if (rn[r] <= tol * bnorm[r] || its[r] >= maxit) act[r] = 0;
        }
    }
    return NRQ_OK;
}

// ---------------------------------------------------------------------------
// Chebyshev iteration.  A_r = 1 + beta_r K is normal and K has its spectrum in
// [0, kmax], kmax = sum_i 2/a_i^2 (K = sum_i (1 - (T_i + T_i^+)/2)/a_i^2 with
// unitary hops T_i), so the spectrum of A_r lies on the segment from 1 to
// 1 + beta_r kmax.  With centre c and half-width d of that segment and
// sigma = c/d, the residual after k steps is bounded by ||r0|| / |T_k(sigma)|.
// ---------------------------------------------------------------------------
static double cheb_kmax()
{
    const geom &g = lat().g;
    return 2.0 / (g.a[1] * g.a[1]) + 2.0 / (g.a[2] * g.a[2]) + 2.0 / (g.a[3] * g.a[3]);
}

// Smallest k <= CHEB_MAX with max_r 1/|T_k(sigma_r)| <= target.
static int cheb_steps(const std::vector<std::complex<double>> &sig, double target)
{
    for (int k = 1; k < CHEB_MAX; ++k) {
        double worst = 0;
        for (const auto &sg : sig) {
            std::complex<double> t0 = 1.0, t1 = sg;
            for (int j = 1; j < k; ++j) {
                const std::complex<double> t2 = 2.0 * sg * t1 - t0;
                t0 = t1;
                t1 = t2;
            }
            worst = std::max(worst, 1.0 / std::abs(t1));
        }
        if (worst <= target) return k;
    }
    return CHEB_MAX;
}

// k Chebyshev steps on A x = b from x (result back in x); xalt is scratch.
template <typename RL, typename T>
static int cheb_run(const cplx<RL> *links, int R, const cplx<T> *b, cplx<T> *x, cplx<T> *xalt,
                    const std::vector<std::complex<double>> &beta, int k, double kmax, const std::vector<int> &act)
{
    std::vector<cplx<double>> tab((size_t)2 * NSLOTS * k);
    for (int r = 0; r < R; ++r) {
        const std::complex<double> c = 1.0 + beta[r] * (kmax / 2), d = beta[r] * (kmax / 2);
        const std::complex<double> sg = c / d, ic = 1.0 / c;
        std::complex<double> rho = 1.0;
        for (int j = 0; j < k; ++j) {
            if (j == 1) rho = 1.0 / (1.0 - 1.0 / (2.0 * sg * sg));
            else if (j > 1) rho = 1.0 / (1.0 - rho / (4.0 * sg * sg));
            tab[(size_t)2 * NSLOTS * j + 2 * r] = {ic.real(), ic.imag()};
            tab[(size_t)2 * NSLOTS * j + 2 * r + 1] = {rho.real(), rho.imag()};
        }
    }
    NRQ_HIP(hipMemcpyAsync(g_ws.cheb, tab.data(), sizeof(cplx<double>) * tab.size(), hipMemcpyHostToDevice, stream()));
    const geom g = lat().g;
    const lf_map mp = make_lf_map(g, 256, 8, 1, (R + RCHUNK - 1) / RCHUNK);
    cplx<T> *cur = x, *other = xalt;
    for (int j = 0; j < k; ++j) {
        hipLaunchKernelGGL((k_cheb<RL, T>), dim3((unsigned)(mp.ngroup * mp.nsb)), dim3(256), 0, stream(), links,
                           (const cplx<T> *)cur, other, b, g_ws.beta, g_ws.cheb + (size_t)2 * NSLOTS * j, g_ws.active, R, g,
                           mp, j == 0 ? 1 : 0);
        std::swap(cur, other);
    }
    NRQ_HIP(hipGetLastError());
    if (cur != x) {
        // odd step count: the result is in xalt; only active slots were iterated
        const size_t fb = sizeof(cplx<T>) * (size_t)3 * g.vol;
        for (int r = 0; r < R; ++r)
            if (act[r]) NRQ_HIP(hipMemcpyAsync((char *)x + fb * r, (const char *)cur + fb * r, fb, hipMemcpyDeviceToDevice, stream()));
    }
    return NRQ_OK;
}

// FP64 true residual res = y - A x for the active slots; ||res||^2 to host.
static int residual64(const nrq_gauge *gf, int R, double *rn2)
{
    if (gf->precision == 64) {
        if (int e = apply_t<double, double>((const cplx<double> *)gauge_links(gf), g_ws.x, g_ws.res, R, nullptr, g_ws.y)) return e;
    } else {
        if (int e = apply_t<float, double>((const cplx<float> *)gauge_links(gf), g_ws.x, g_ws.res, R, nullptr, g_ws.y)) return e;
    }
    return norms2_t<double>(g_ws.res, R, rn2);
}

static int cheb_solve(const nrq_gauge *gf, int R, const std::vector<std::complex<double>> &beta, const double *bn,
                      const krylov_params &p, std::vector<int> &act, std::vector<double> &rel, std::vector<int64_t> &its,
                      int64_t &cycles)
{
    const int64_t n = 3 * g_ws.vol;
    const double kmax = cheb_kmax();
    std::vector<std::complex<double>> sig(R);
    for (int r = 0; r < R; ++r) {
        const std::complex<double> d = beta[r] * (kmax / 2);
        sig[r] = std::abs(d) > 0 ? (1.0 + d) / d : std::complex<double>(1e30, 0.0);
    }
    const bool mixed = p.kind == KRYLOV_CHEB_MIXED;
    const cplx<float> *l32 = nullptr;
    if (mixed) {
        if (gf->precision == 64) {
            hipLaunchKernelGGL(k_links_to_f32, dim3(4096), dim3(256), 0, stream(), (const cplx<double> *)gauge_links(gf),
                               g_ws.links32, 27 * g_ws.vol);
            l32 = g_ws.links32;
        } else {
            l32 = (const cplx<float> *)gauge_links(gf);
        }
    }
    double rn2[NSLOTS];
    for (int outer = 0; outer < 8; ++outer) {
        if (int e = upload_active(act)) return e;
        if (int e = residual64(gf, R, rn2)) return e;
        double need = 1.0;
        bool any = false;
        for (int r = 0; r < R; ++r) {
            if (!act[r]) continue;
            const double rr = std::sqrt(std::max(rn2[r], 0.0));
            rel[r] = rr / bn[r];
            if (rr <= p.tol * bn[r]) {
                act[r] = 0;
                continue;
            }
            any = true;
            need = std::min(need, 0.3 * p.tol * bn[r] / rr);
        }
        if (!any) return NRQ_OK;
        ++cycles;
        if (int e = upload_active(act)) return e;
        if (!mixed) {
            const int k = cheb_steps(sig, need);
            if (gf->precision == 64) {
                if (int e = cheb_run<double, double>((const cplx<double> *)gauge_links(gf), R, g_ws.y, g_ws.x, g_ws.t, beta, k, kmax, act))
                    return e;
            } else {
                if (int e = cheb_run<float, double>((const cplx<float> *)gauge_links(gf), R, g_ws.y, g_ws.x, g_ws.t, beta, k, kmax, act))
                    return e;
            }
            for (int r = 0; r < R; ++r) its[r] += act[r] ? k : 0;
        } else {
            // correction A d = r / ||r|| in FP32 to the accuracy FP32 allows
            const int k = cheb_steps(sig, std::max(need, 3e-7));
            double *rnorm2 = g_ws.nrm + 32;
            NRQ_HIP(hipMemcpyAsync(rnorm2, g_ws.nrm, sizeof(double) * R, hipMemcpyDeviceToDevice, stream()));
            hipLaunchKernelGGL(k_demote, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), (const cplx<double> *)g_ws.res, g_ws.y32, n,
                               (const double *)rnorm2, g_ws.active);
            NRQ_HIP(hipMemsetAsync(g_ws.x32, 0, sizeof(cplx<float>) * (size_t)n * R, stream()));
            if (int e = cheb_run<float, float>(l32, R, g_ws.y32, g_ws.x32, g_ws.t32, beta, k, kmax, act)) return e;
            hipLaunchKernelGGL(k_promote_add, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), g_ws.x, (const cplx<float> *)g_ws.x32, n,
                               (const double *)rnorm2, g_ws.active);
            NRQ_HIP(hipGetLastError());
            for (int r = 0; r < R; ++r) its[r] += act[r] ? k : 0;
        }
    }
    return NRQ_OK; // slots still active fall back to GMRES in the caller
}

int krylov_solve(const nrq_gauge *gf, int R, const cplx<double> *beta, const cplx<double> *alpha,
                 const cplx<double> *prev, const double *y_norm2, const krylov_params &p, nrq_solve_stats *st)
{
    const int64_t n = 3 * g_ws.vol;
    std::vector<double> bn(R);
    for (int r = 0; r < R; ++r) bn[r] = std::sqrt(y_norm2[r]);
    std::vector<int64_t> its(R, 0);
    std::vector<double> rel(R, 0.0);
    int64_t cycles = 0;
    std::vector<int> act(R, 1);
    if (int e = upload_beta(beta, R)) return e;
    int status = NRQ_OK;

    if (p.kind == KRYLOV_GMRES_FP64) {
        if (gf->precision == 64)
            status = gmres_core<double, double>((const cplx<double> *)gauge_links(gf), R, g_ws.y, g_ws.x,
                                                (cplx<double> *)g_ws.basis, g_ws.t, g_ws.res, true, bn.data(), act, p.tol,
                                                p.restart, p.maxit, its, rel.data(), cycles);
        else
            status = gmres_core<double, float>((const cplx<float> *)gauge_links(gf), R, g_ws.y, g_ws.x,
                                               (cplx<double> *)g_ws.basis, g_ws.t, g_ws.res, true, bn.data(), act, p.tol,
                                               p.restart, p.maxit, its, rel.data(), cycles);
        if (status) return status;
    } else if (p.kind == KRYLOV_CHEB_FP64 || p.kind == KRYLOV_CHEB_MIXED) {
        std::vector<std::complex<double>> bz(R);
        for (int r = 0; r < R; ++r) bz[r] = std::complex<double>(beta[r].re, beta[r].im);
        if (int e = upload_active(act)) return e;
        if (p.extrapolate && prev && alpha) {
            std::vector<cplx<double>> coef(R);
            for (int r = 0; r < R; ++r) {
                const std::complex<double> a(alpha[r].re, alpha[r].im);
                const std::complex<double> c = std::abs(a) > 0 ? (a - bz[r]) / a : std::complex<double>(0.0);
                coef[r] = {c.real(), c.imag()};
            }
            NRQ_HIP(hipMemcpyAsync(g_ws.coef, coef.data(), sizeof(cplx<double>) * R, hipMemcpyHostToDevice, stream()));
            hipLaunchKernelGGL(k_guess, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), g_ws.x, prev, g_ws.y, g_ws.coef, n, g_ws.active);
        } else {
            NRQ_HIP(hipMemsetAsync(g_ws.x, 0, sizeof(cplx<double>) * (size_t)n * R, stream()));
        }
        if (int e = cheb_solve(gf, R, bz, bn.data(), p, act, rel, its, cycles)) return e;
        bool left = false;
        for (int r = 0; r < R; ++r) left = left || act[r];
        if (left) {
            // spectrum bound violated (e.g. far from unitary links): GMRES from here
            if (gf->precision == 64)
                status = gmres_core<double, double>((const cplx<double> *)gauge_links(gf), R, g_ws.y, g_ws.x,
                                                    (cplx<double> *)g_ws.basis, g_ws.t, g_ws.res, false, bn.data(), act,
                                                    p.tol, p.restart, p.maxit, its, rel.data(), cycles);
            else
                status = gmres_core<double, float>((const cplx<float> *)gauge_links(gf), R, g_ws.y, g_ws.x,
                                                   (cplx<double> *)g_ws.basis, g_ws.t, g_ws.res, false, bn.data(), act,
                                                   p.tol, p.restart, p.maxit, its, rel.data(), cycles);
            if (status) return status;
        }
    } else {
        // FP32 links for the corrections
        const cplx<float> *l32;
        if (gf->precision == 64) {
            hipLaunchKernelGGL(k_links_to_f32, dim3(4096), dim3(256), 0, stream(), (const cplx<double> *)gauge_links(gf),
                               g_ws.links32, 27 * g_ws.vol);
            l32 = g_ws.links32;
        } else {
            l32 = (const cplx<float> *)gauge_links(gf);
        }
        // initial guess
        if (int e = upload_active(act)) return e;
        if (p.extrapolate && prev && alpha) {
            std::vector<cplx<double>> coef(R);
            for (int r = 0; r < R; ++r) {
                const std::complex<double> a(alpha[r].re, alpha[r].im), b(beta[r].re, beta[r].im);
                const std::complex<double> c = std::abs(a) > 0 ? (a - b) / a : std::complex<double>(0.0);
                coef[r] = {c.real(), c.imag()};
            }
            NRQ_HIP(hipMemcpyAsync(g_ws.coef, coef.data(), sizeof(cplx<double>) * R, hipMemcpyHostToDevice, stream()));
            hipLaunchKernelGGL(k_guess, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), g_ws.x, prev, g_ws.y, g_ws.coef, n, g_ws.active);
        } else {
            NRQ_HIP(hipMemsetAsync(g_ws.x, 0, sizeof(cplx<double>) * (size_t)n * R, stream()));
        }
        std::vector<double> one(R, 1.0);
        for (int outer = 0; outer < 40; ++outer) {
            // FP64 true residual r = y - A x
            if (int e = upload_active(act)) return e;
            if (int e = apply64(gf, g_ws.x, g_ws.t, R)) return e;
            hipLaunchKernelGGL(k_residual<double>, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), (const cplx<double> *)g_ws.y,
                               (const cplx<double> *)g_ws.t, g_ws.res, n, g_ws.active, g_ws.rpart);
            hipLaunchKernelGGL(k_fold_r, dim3(1), dim3(64), 0, stream(), g_ws.rpart, VBLOCKS, R, g_ws.active, g_ws.nrm);
            NRQ_HIP(hipGetLastError());
            NRQ_HIP(hipMemcpyAsync(g_ws.nbuf, g_ws.nrm, sizeof(double) * R, hipMemcpyDeviceToHost, stream()));
            NRQ_HIP(hipStreamSynchronize(stream()));
            bool any = false;
            for (int r = 0; r < R; ++r) {
                if (!act[r]) continue;
                const double rr = std::sqrt(std::max(g_ws.nbuf[r], 0.0));
                rel[r] = rr / bn[r];
                if (rr <= p.tol * bn[r] || its[r] >= p.maxit) act[r] = 0;
                any = any || act[r];
            }
            if (!any) break;
            ++cycles;
            // keep ||r||^2 of this refinement step in nrm[32..) (the inner
            // solve reuses nrm[0..R))
            double *rnorm2 = g_ws.nrm + 32;
            NRQ_HIP(hipMemcpyAsync(rnorm2, g_ws.nrm, sizeof(double) * R, hipMemcpyDeviceToDevice, stream()));
            // normalised FP32 correction problem A d = r / ||r||
            if (int e = upload_active(act)) return e;
            hipLaunchKernelGGL(k_demote, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), (const cplx<double> *)g_ws.res, g_ws.y32, n,
                               (const double *)rnorm2, g_ws.active);
            std::vector<double> inner_rel(R, 0.0);
            int64_t inner_cycles = 0;
            if (int e = gmres_core<float, float>(l32, R, g_ws.y32, g_ws.x32, (cplx<float> *)g_ws.basis, g_ws.t32, g_ws.res32, true,
                                                 one.data(), act, std::max(p.inner_tol, std::max(p.tol / 1e-3, 2e-7)), p.restart,
                                                 p.maxit, its, inner_rel.data(), inner_cycles))
                return e;
            // gmres_core left the slot masks of the inner solve on the device;
            // the refinement updates every slot that is still active here
            if (int e = upload_active(act)) return e;
            hipLaunchKernelGGL(k_promote_add, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), g_ws.x, (const cplx<float> *)g_ws.x32, n,
                               (const double *)rnorm2, g_ws.active);
            NRQ_HIP(hipGetLastError());
        }
        for (int r = 0; r < R; ++r)
            if (act[r]) status = NRQ_ERR_NOCONV;
    }
    double worst = 0;
    int64_t total = 0;
    for (int r = 0; r < R; ++r) {
        worst = std::max(worst, rel[r]);
        total += its[r];
        if (rel[r] > p.tol) status = NRQ_ERR_NOCONV;
    }
    if (st) {
        st->iterations += total;
        st->restarts += cycles;
        st->max_true_relres = std::max(st->max_true_relres, worst);
    }
    if (status == NRQ_ERR_NOCONV)
        return fail(NRQ_ERR_NOCONV, "CN: Krylov solver stopped (maxit=%lld) with true relative residual %.3e > %.1e",
                    (long long)p.maxit, worst, p.tol);
    return status;
}

} // namespace nrq
