// nrq_krylov.hip -- batched solvers for the NRQCD Crank-Nicolson systems
// (replaces PETSc KSP, nrqcd.f90:786-804, 1103-1105, 1123, 1161).
//
// A CN step of R colour-vector fields psi_r solves
//   (1 + beta_r K) x_r = y_r,   y_r = (1 + alpha_r K) psi_r
// for all r at once; every stencil launch streams the links once per chunk of
// NC slots.  Solvers:
//  * KRYLOV_GMRES_FP64: GMRES(m), classical Gram-Schmidt, FP64, zero guess
//    (PETSc's defaults).
//  * KRYLOV_GMRES_MIXED: FP64 iterative refinement around FP32 GMRES(m)
//    corrections.
//  * KRYLOV_FGMRES_CNN: flexible GMRES(m) in FP64 whose right preconditioner
//    is z = v - f1 + f2 + N(v, f1, f2) with f1 = beta K v, f2 = beta K f1:
//    the second-order Neumann polynomial of A^-1 corrected by the learned
//    network of nrq_cnn.hip, which sees each direction with its two
//    covariant features (K is the covariant Laplacian, so the inputs
//    transform like v under gauge transformations, and the beta scaling
//    makes them mass-independent).
//  * KRYLOV_FGMRES_CNN_MIXED: FP64 iterative refinement around FP32 FGMRES(m)
//    corrections with the same network.
//  * KRYLOV_CHEB_FP64 / KRYLOV_CHEB_MIXED: Chebyshev iteration on the known
//    spectral segment of A_r, in FP64 or as FP32 corrections inside FP64
//    iterative refinement.  Step counts are planned from a priori bounds, so
//    the usual CN step needs one host synchronisation; the FP64 true residual
//    then confirms convergence (and more steps, or GMRES, follow if not).
// In every case ||y - A x|| <= tol ||y|| with the FP64 operator decides
// convergence: precision and solver change the work, never the answer.
#include "nrq_krylov.h"
#include "nrq_cnn.h"
#include "nrq_map.h"
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace nrq {

const void *gauge_links(const nrq_gauge *h);

constexpr int CHEB_MAX = 64;  // Chebyshev steps per coefficient table
constexpr int VBLOCKS = 256;  // blocks per slot in the vector kernels
constexpr int VTHREADS = 256;
constexpr int SBS = 256;      // threads per block in the stencil kernels (site-block maps)
constexpr int SBS_WAVE = 512; // at most, with one work item per wavefront (wave-item map)
#ifndef CN_WAVES
#define CN_WAVES 4 // wavefronts per SIMD the stencil kernels' registers must allow
#endif
constexpr double FP32_CORR_ERR = 1e-6; // relative error assumed for an FP32 correction

// ---------------------------------------------------------------------------
// device helpers
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

template <typename T> __device__ __forceinline__ double nrm2(cplx<T> a) { return (double)a.re * a.re + (double)a.im * a.im; }

template <typename T> __device__ __forceinline__ cplx<T> to(cplx<double> a) { return {(T)a.re, (T)a.im}; }

__device__ __forceinline__ double g_dsum(const geom &g)
{
    return 1.0 / (g.a[1] * g.a[1]) + 1.0 / (g.a[2] * g.a[2]) + 1.0 / (g.a[3] * g.a[3]);
}

// Base pointers of the slot fields of one launch (3V complex numbers each,
// colour-major); a slot's field may live in the workspace or in the heavy
// field's own storage.  Passed to the kernels by value.
template <typename T> struct slots {
    cplx<T> *p[NSLOTS];
};

// Per-slot fields of one chunk, read in the compute type T (NT: streaming
// loads, for values no other thread reads).  The slot pointers are the same
// for a whole block, so every load is a scalar base plus a 32-bit offset.
template <int NC, typename T, typename S, bool NT = false> struct ld_fields {
    const cplx<S> *f[NC];
    __device__ __forceinline__ cplx<T> operator()(int q, int c, int64_t vol, uint32_t site) const
    {
        const cplx<S> *p = at32(f[q] + (int64_t)c * vol, site);
        const cplx<S> v = NT ? load_nt(p) : *p;
        return {(T)v.re, (T)v.im};
    }
};

// hop_q = sum_i (U_i(x) f_q(x+i) + U_i(x-i)^+ f_q(x-i)) / a_i^2 at x = s for
// the slots q < NC of one chunk; with dsum = sum_i 1/a_i^2 the kinetic
// operator of nrqcd.f90:1210-1242 (without its coefficient) is
//   K f(s) = dsum f(s) - hop / 2
//          = sum_i [f(s) - (U_i(s) f(s+i) + U_i(s-i)^+ f(s-i))/2] / a_i^2.
// ld(q, c, V, site) returns colour c of slot q's field at site; it
// must be readable for every q (a chunk's unused slots read a used slot's
// field, and the caller drops their results).  The kernels are bound by
// memory latency, so each of the six hops issues its 9 link loads and all
// 3 NC field loads at once, with no branch between them: one round trip per
// hop.  Only rows 1 and 2 of each link are read and row 3 is rebuilt as
// conj(row1 x row2), exact for SU(3) up to the links' rounding, a third fewer
// link loads, which are half of these kernels' memory time: for FP64 links
// (rebuilt to FP64 rounding) and in FP32 arithmetic on FP32 links (the
// correction solves, whose results the FP64 residual checks).  FP64
// arithmetic on FP32-stored links reads whole links, so that the FP64
// residual is that of the stored operator and not of its SU(3) projection.
template <int NC, typename T, typename RL, typename LD>
__device__ __forceinline__ void hop_chunk(const cplx<RL> *__restrict__ u, const LD &ld, const lsite &p, const geom &g,
                                          cplx<T> (&hop)[NC][3])
{
#pragma unroll
    for (int q = 0; q < NC; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) hop[q][c] = {T(0), T(0)};
    // a loop, not unrolled: unrolled, the compiler hoists the loads of all six
    // hops to the top and spills them
#pragma unroll 1
    for (int h = 0; h < 6; ++h) {
        const int i = h >> 1;
        const bool fwd = (h & 1) == 0;
        const T ia2 = T(1.0 / (g.a[i + 1] * g.a[i + 1]));
        // forward: U_i(s) and f(s+i); backward: U_i(s-i) and f(s-i)
        const int64_t sl = fwd ? p.s : site_dn(g, p, i).s;
        const int64_t sf = fwd ? site_up(g, p, i).s : sl;
        constexpr bool REC = sizeof(RL) == 8 || sizeof(T) == 4;
        constexpr int NE = REC ? 6 : 9;
        cplx<RL> lk[9];
        cplx<T> f[NC][3];
#pragma unroll
        for (int e = 0; e < NE; ++e) lk[e] = *at32(u + ((int64_t)i * 9 + e) * g.vol, (uint32_t)sl);
#pragma unroll
        for (int q = 0; q < NC; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) f[q][c] = ld(q, c, g.vol, (uint32_t)sf);
        // hop += (U f) / a_i^2, the product in the compute type
        cplx<T> w[9];
#pragma unroll
        for (int e = 0; e < NE; ++e) w[e] = {(T)lk[e].re, (T)lk[e].im};
        if (REC)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                cplx<T> v = {T(0), T(0)};
                mac_cc(v, w[c1], w[3 + c2]);
                const cplx<T> a = {-w[c2].re, -w[c2].im};
                mac_cc(v, a, w[3 + c1]);
                w[6 + c] = v;
            }
#pragma unroll
        for (int q = 0; q < NC; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                cplx<T> t = {T(0), T(0)};
                if (fwd) {
                    mac(t, w[3 * c], f[q][0]);
                    mac(t, w[3 * c + 1], f[q][1]);
                    mac(t, w[3 * c + 2], f[q][2]);
                } else {
                    mac_cn(t, w[c], f[q][0]);
                    mac_cn(t, w[3 + c], f[q][1]);
                    mac_cn(t, w[6 + c], f[q][2]);
                }
                hop[q][c].re = fmx(ia2, t.re, hop[q][c].re);
                hop[q][c].im = fmx(ia2, t.im, hop[q][c].im);
            }
    }
}

// The chunk's own values of slot fields: v[q][c] = ld(q, c, V, s), all
// issued before any store of the caller.
template <int NC, typename T, typename LD>
__device__ __forceinline__ void own_chunk(const LD &ld, int64_t s, int64_t vol, cplx<T> (&v)[NC][3])
{
#pragma unroll
    for (int q = 0; q < NC; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) v[q][c] = ld(q, c, vol, (uint32_t)s);
}

// Per-slot sums of v[q] over the block (wave = false; every thread of the
// block must call this) or over the wavefront (wave = true: the wave-item
// map, one chunk per wavefront).  The sum of slot r0 + q < nf goes to
// partial[(r0 + q) * nsb + t].
template <int NC>
__device__ __forceinline__ void chunk_partials(double (&v)[NC], int r0, int nf, int64_t t, int64_t nsb,
                                               double *__restrict__ partial, bool wave)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (wave) {
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const double x = wsum(v[q]);
            if (lane == 0 && r0 + q < nf) partial[(int64_t)(r0 + q) * nsb + t] = x;
        }
        return;
    }
    __shared__ double sh[NC][SBS / 64];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        const double x = wsum(v[q]);
        if (lane == 0) sh[q][w] = x;
    }
    __syncthreads();
    if (threadIdx.x < NC && r0 + (int)threadIdx.x < nf) {
        const int q = threadIdx.x;
        double x = 0;
        for (int k = 0; k < SBS / 64; ++k) x += sh[q][k];
        partial[(int64_t)(r0 + q) * nsb + t] = x;
    }
}

// ---------------------------------------------------------------------------
// stencil kernels (grid: lf_map with one work item per chunk of NC slots)
// ---------------------------------------------------------------------------

// The slots of chunk r0 that a stencil kernel computes: on[q] for slot
// r0 + q, and every load pointer of an unused slot aimed at the first used
// one (same addresses, so its loads hit the cache).  False when no slot of
// the chunk is used.
template <int NC>
__device__ __forceinline__ bool chunk_slots(int r0, int nf, const int *__restrict__ active, bool (&on)[NC], int (&rs)[NC])
{
    int first = -1;
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        on[q] = r0 + q < nf && (!active || active[r0 + q]);
        if (on[q] && first < 0) first = r0 + q;
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) rs[q] = on[q] ? r0 + q : first;
    return first >= 0;
}

// out_r = in_r + beta_r K in_r for the active slots; with gco (not null)
// instead out_r = -beta_r gco_r K in_r.
template <int NC, typename RL, typename T>
__global__ __launch_bounds__(SBS_WAVE, CN_WAVES) void k_stencil(const cplx<RL> *__restrict__ u, slots<T> in, slots<T> out,
                                                 const cplx<double> *__restrict__ beta, const int *__restrict__ active,
                                                 int nf, geom g, lf_map mp, const cplx<double> *__restrict__ gco)
{
    lsite p;
    int chunk;
    if (!lf_site(mp, g, p, chunk)) return;
    const int64_t s = p.s;
    const int r0 = chunk * NC;
    bool on[NC];
    int rs[NC];
    if (!chunk_slots<NC>(r0, nf, active, on, rs)) return;
    ld_fields<NC, T, T> ld;
#pragma unroll
    for (int q = 0; q < NC; ++q) ld.f[q] = in.p[rs[q]];
    cplx<T> hop[NC][3];
    hop_chunk<NC, T>(u, ld, p, g, hop);
    cplx<T> v0[NC][3];
    own_chunk<NC, T>(ld, s, g.vol, v0);
    const T dsum = T(g_dsum(g));
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        if (!on[q]) continue;
        cplx<T> *o = out.p[r0 + q];
        cplx<T> cf;
        if (gco) {
            const cplx<double> b = beta[r0 + q], gc = gco[r0 + q];
            cf = to<T>(cplx<double>{-(b.re * gc.re - b.im * gc.im), -(b.re * gc.im + b.im * gc.re)});
        } else {
            cf = to<T>(beta[r0 + q]);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const cplx<T> kf = dsum * v0[q][c] - T(0.5) * hop[q][c];
            store_nt(at32(o + (int64_t)c * g.vol, (uint32_t)s), gco ? cf * kf : v0[q][c] + cf * kf);
        }
    }
}

// Right-hand side and initial guess of a CN step for slot r:
//   y_r = psi_r + alpha_r K psi_r                     (S psi, nrqcd.f90:1112)
//   x_r = psi_r + g_r K psi_r  (guess != 0)  or  x_r = 0
// with g_r = alpha_r - beta_r the first-order expansion of
// (1 + beta K)^-1 (1 + alpha K).  Partial ||y_r||^2 per site block.  With wk,
// K psi_r also goes to kout in FP32: the guess's residual is exactly
// y - (1 + beta K) x = -beta g K^2 psi.
template <int NC, typename RL>
__global__ __launch_bounds__(SBS_WAVE, CN_WAVES) void k_cn_rhs(const cplx<RL> *__restrict__ u, slots<double> psi, slots<double> y,
                                                slots<double> x, const cplx<double> *__restrict__ alpha,
                                                const cplx<double> *__restrict__ gco, int guess, int nf, geom g, lf_map mp,
                                                double *__restrict__ partial, slots<float> kout, int wk)
{
    lsite p = {0, 0, 0, 0};
    int64_t t = 0;
    int chunk;
    const bool valid = lf_site(mp, g, p, chunk, t);
    const int64_t s = p.s;
    const int r0 = chunk * NC;
    double nrm[NC] = {};
    bool on[NC];
    int rs[NC];
    if (valid && chunk_slots<NC>(r0, nf, nullptr, on, rs)) {
        ld_fields<NC, double, double> ld;
#pragma unroll
        for (int q = 0; q < NC; ++q) ld.f[q] = psi.p[rs[q]];
        cplx<double> hop[NC][3];
        hop_chunk<NC, double>(u, ld, p, g, hop);
        cplx<double> v0[NC][3];
        own_chunk<NC, double>(ld, s, g.vol, v0);
        const double dsum = g_dsum(g);
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            if (!on[q]) continue;
            const int r = r0 + q;
            const cplx<double> a = alpha[r], gc = gco[r];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int64_t cv = (int64_t)c * g.vol;
                const cplx<double> kf = dsum * v0[q][c] - 0.5 * hop[q][c];
                const cplx<double> yv = v0[q][c] + a * kf;
                store_nt(at32(y.p[r] + cv, (uint32_t)s), yv);
                store_nt(at32(x.p[r] + cv, (uint32_t)s), guess ? v0[q][c] + gc * kf : cplx<double>{0.0, 0.0});
                if (wk) store_nt(at32(kout.p[r] + cv, (uint32_t)s), cplx<float>{(float)kf.re, (float)kf.im});
                nrm[q] += nrm2(yv);
            }
        }
    }
    chunk_partials<NC>(nrm, r0, nf, t, mp.nsb, partial, mp.mode >= 3);
}

// FP64 true residual r_r = y_r - (1 + beta_r K) x_r, stored to res as type
// TR when WRITE_RES; partial ||r_r||^2 per site block.
template <int NC, typename RL, typename TR, bool WRITE_RES>
__global__ __launch_bounds__(SBS_WAVE, CN_WAVES) void k_cn_resid(const cplx<RL> *__restrict__ u, slots<double> x, slots<double> y,
                                                  slots<TR> res, const cplx<double> *__restrict__ beta, int nf, geom g,
                                                  lf_map mp, double *__restrict__ partial)
{
    lsite p = {0, 0, 0, 0};
    int64_t t = 0;
    int chunk;
    const bool valid = lf_site(mp, g, p, chunk, t);
    const int64_t s = p.s;
    const int r0 = chunk * NC;
    double nrm[NC] = {};
    bool on[NC];
    int rs[NC];
    if (valid && chunk_slots<NC>(r0, nf, nullptr, on, rs)) {
        ld_fields<NC, double, double> ld;
        ld_fields<NC, double, double, true> ly;
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            ld.f[q] = x.p[rs[q]];
            ly.f[q] = y.p[rs[q]];
        }
        cplx<double> hop[NC][3];
        hop_chunk<NC, double>(u, ld, p, g, hop);
        cplx<double> v0[NC][3], yv[NC][3];
        own_chunk<NC, double>(ld, s, g.vol, v0);
        own_chunk<NC, double>(ly, s, g.vol, yv);
        const double dsum = g_dsum(g);
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            if (!on[q]) continue;
            const int r = r0 + q;
            const cplx<double> bt = beta[r];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const cplx<double> rr = yv[q][c] - (v0[q][c] + bt * (dsum * v0[q][c] - 0.5 * hop[q][c]));
                if (WRITE_RES) store_nt(at32(res.p[r] + (int64_t)c * g.vol, (uint32_t)s), cplx<TR>{(TR)rr.re, (TR)rr.im});
                nrm[q] += nrm2(rr);
            }
        }
    }
    chunk_partials<NC>(nrm, r0, nf, t, mp.nsb, partial, mp.mode >= 3);
}

// One Chebyshev step on the segment [1, 1 + beta_r kmax] of the spectrum of
// A_r = 1 + beta_r K, fused with the operator application:
//   x_{k+1} = rho_r (x_k + (b_r - A_r x_k) / c_r) + (1 - rho_r) x_{k-1}
// c_r = 1 + beta_r kmax / 2 is the segment centre.  mode 0: as above, x_{k-1}
// read from xprev; 1: first step (rho = 1, no x_{k-1}); 2: from a zero start,
// x_k = b / c implicit (xk is b) and x_{k-1} = 0; 3: x_{k-1} = b / c implicit.
// x_{k+1} goes to xnext, or with fin to fout = fin + x_{k+1} in FP64 (the
// last step of a correction solve adds the correction to the FP64 iterate).
// xprev and xnext may be the same fields: each thread reads its own values of
// x_{k-1} before it writes them.
template <int NC, typename RL, typename T>
__global__ __launch_bounds__(SBS_WAVE, CN_WAVES) void k_cheb(const cplx<RL> *__restrict__ u, slots<T> xk, slots<T> xprev, slots<T> xnext,
                                              slots<T> b, const cplx<double> *__restrict__ beta,
                                              const cplx<double> *__restrict__ coef, const int *__restrict__ active, int nf,
                                              geom g, lf_map mp, int mode, int fin, slots<double> fin_in,
                                              slots<double> fin_out)
{
    lsite p;
    int chunk;
    if (!lf_site(mp, g, p, chunk)) return;
    const int64_t s = p.s;
    const int r0 = chunk * NC;
    bool on[NC];
    int rs[NC];
    if (!chunk_slots<NC>(r0, nf, active, on, rs)) return;
    ld_fields<NC, T, T> ld;
    ld_fields<NC, T, T, true> lb, lp;
    ld_fields<NC, double, double, true> lf;
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        ld.f[q] = xk.p[rs[q]];
        lb.f[q] = b.p[rs[q]];
        lp.f[q] = xprev.p[rs[q]];
        lf.f[q] = fin_in.p[rs[q]];
    }
    cplx<T> hop[NC][3];
    hop_chunk<NC, T>(u, ld, p, g, hop);
    cplx<T> v0[NC][3], bs[NC][3];
    own_chunk<NC, T>(ld, s, g.vol, v0);
    if (mode != 2) own_chunk<NC, T>(lb, s, g.vol, bs);
    const T dsum = T(g_dsum(g));
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        if (!on[q]) continue;
        const int r = r0 + q;
        const cplx<T> bt = to<T>(beta[r]);
        const cplx<T> ic = to<T>(coef[2 * r]); // 1 / c_r
        const cplx<T> rho = to<T>(coef[2 * r + 1]);
        const cplx<T> one_m_rho = {T(1) - rho.re, -rho.im};
        // x_{k-1} and the FP64 iterate per slot (xprev may be xnext)
        cplx<T> xp[3];
        cplx<double> xo[3];
        if (mode == 0)
#pragma unroll
            for (int c = 0; c < 3; ++c) xp[c] = lp(q, c, g.vol, (uint32_t)s);
        if (fin)
#pragma unroll
            for (int c = 0; c < 3; ++c) xo[c] = lf(q, c, g.vol, (uint32_t)s);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int64_t cv = (int64_t)c * g.vol;
            const cplx<T> kf = dsum * v0[q][c] - T(0.5) * hop[q][c];
            const cplx<T> bq = mode == 2 ? v0[q][c] : bs[q][c];
            const cplx<T> x0 = mode == 2 ? ic * v0[q][c] : v0[q][c];
            const cplx<T> ax = mode == 2 ? ic * kf : kf;
            cplx<T> nx = rho * (x0 + ic * (bq - (x0 + bt * ax)));
            if (mode == 3) nx = nx + one_m_rho * (ic * bq);
            else if (mode == 0) nx = nx + one_m_rho * xp[c];
            if (fin)
                store_nt(at32(fin_out.p[r] + cv, (uint32_t)s),
                         cplx<double>{xo[c].re + (double)nx.re, xo[c].im + (double)nx.im});
            else
                store_nt(at32(xnext.p[r] + cv, (uint32_t)s), nx);
        }
    }
}

// ---------------------------------------------------------------------------
// vector kernels (grid: VBLOCKS x slots)
// ---------------------------------------------------------------------------

// partial[r][i][blk] = this block's share of <V_i,r, w_r>, i < nv (FP64 sums).
// The basis vectors go in groups of MDOT_G, each group in one pass that reads
// w_r once (PETSc's VecMDot); per thread and per block the sums run in the
// same order for every grouping.
constexpr int MDOT_G = 8;
template <typename T>
__global__ __launch_bounds__(VTHREADS) void k_mdot(const cplx<T> *__restrict__ basis, int nv, const cplx<T> *__restrict__ w,
                                                   int64_t n, int nf, const int *__restrict__ active,
                                                   cplx<double> *__restrict__ partial)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    __shared__ double sh[VTHREADS / 64][2 * MDOT_G];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const cplx<T> *wr = w + (int64_t)r * n;
    for (int i0 = 0; i0 < nv; i0 += MDOT_G) {
        const int gn = nv - i0 < MDOT_G ? nv - i0 : MDOT_G;
        double re[MDOT_G], im[MDOT_G];
#pragma unroll
        for (int j = 0; j < MDOT_G; ++j) re[j] = im[j] = 0;
        for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
            const cplx<T> b = wr[k];
#pragma unroll
            for (int j = 0; j < MDOT_G; ++j) {
                if (j >= gn) break;
                const cplx<T> a = basis[((int64_t)(i0 + j) * nf + r) * n + k];
                re[j] += (double)a.re * b.re + (double)a.im * b.im;
                im[j] += (double)a.re * b.im - (double)a.im * b.re;
            }
        }
        __syncthreads(); // sh is free (previous group read)
#pragma unroll
        for (int j = 0; j < MDOT_G; ++j) {
            const double x = wsum(re[j]), y = wsum(im[j]);
            if (lane == 0) {
                sh[wv][2 * j] = x;
                sh[wv][2 * j + 1] = y;
            }
        }
        __syncthreads();
        if ((int)threadIdx.x < 2 * gn) {
            double t = 0;
            for (int k = 0; k < VTHREADS / 64; ++k) t += sh[k][threadIdx.x];
            const int j = threadIdx.x >> 1;
            double *o = (double *)&partial[((int64_t)r * nv + i0 + j) * gridDim.x + blockIdx.x];
            o[threadIdx.x & 1] = t;
        }
    }
}

// out[t] = sum over the nblk block partials of t = r * count + i, one
// wavefront per t (lane l sums partials l, l + 64, ..., then the wavefront
// sum): a fixed order, and 64 loads in flight instead of one.
__global__ __launch_bounds__(256) void k_fold_c(const cplx<double> *__restrict__ partial, int nblk, int count, int nf,
                                                const int *__restrict__ active, cplx<double> *__restrict__ out)
{
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (t >= count * nf) return;
    if (!active[t / count]) return;
    double re = 0, im = 0;
    for (int b = lane; b < nblk; b += 64) {
        const cplx<double> v = partial[(int64_t)t * nblk + b];
        re += v.re;
        im += v.im;
    }
    re = wsum(re);
    im = wsum(im);
    if (lane == 0) out[t] = {re, im};
}

// out[r] = sum_b partial[r * nblk + b] for the slot r = blockIdx.x if
// active[r] (active may be null); fixed summation order.  FOLD_THREADS
// threads with four independent sums each keep enough loads in flight for
// the V / 64 partials of the wave-item map.
constexpr int FOLD_THREADS = 1024;
__global__ __launch_bounds__(FOLD_THREADS) void k_fold_slots(const double *__restrict__ partial, int64_t nblk,
                                                             const int *__restrict__ active, double *__restrict__ out)
{
    const int r = blockIdx.x;
    if (active && !active[r]) return;
    __shared__ double sh[FOLD_THREADS / 64];
    const double *p = partial + (int64_t)r * nblk;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    constexpr int64_t STEP = 4 * FOLD_THREADS;
    for (int64_t b = threadIdx.x; b < nblk; b += STEP) {
        s0 += p[b];
        if (b + FOLD_THREADS < nblk) s1 += p[b + FOLD_THREADS];
        if (b + 2 * FOLD_THREADS < nblk) s2 += p[b + 2 * FOLD_THREADS];
        if (b + 3 * FOLD_THREADS < nblk) s3 += p[b + 3 * FOLD_THREADS];
    }
    double s = block_sum((s0 + s1) + (s2 + s3), sh);
    if (threadIdx.x == 0) out[r] = s;
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
        nrm += nrm2(st);
    }
    nrm = block_sum(nrm, sh);
    if (threadIdx.x == 0) partial[(int64_t)r * gridDim.x + blockIdx.x] = nrm;
}

// dst_r = src_r / sqrt(nrm2[r]) (zero norm leaves zeros)
template <typename T>
__global__ void k_scale(const cplx<T> *__restrict__ src, cplx<T> *__restrict__ dst, int64_t n,
                        const double *__restrict__ nrm2v, const int *__restrict__ active)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    const double nn = nrm2v[r];
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
        nrm += nrm2(a);
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
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        nrm += nrm2(x[(int64_t)r * n + k]);
    nrm = block_sum(nrm, sh);
    if (threadIdx.x == 0) partial[(int64_t)r * gridDim.x + blockIdx.x] = nrm;
}

// dst32_r = src64_r / sqrt(nrm2[r])  (normalised FP32 right-hand side)
__global__ void k_demote(const cplx<double> *__restrict__ src, cplx<float> *__restrict__ dst, int64_t n,
                         const double *__restrict__ nrm2v, const int *__restrict__ active)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    const double inv = nrm2v[r] > 0 ? 1.0 / sqrt(nrm2v[r]) : 0.0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const cplx<double> a = src[(int64_t)r * n + k];
        dst[(int64_t)r * n + k] = {(float)(a.re * inv), (float)(a.im * inv)};
    }
}

// x64_r += sqrt(nrm2[r]) d32_r
__global__ void k_promote_add(cplx<double> *__restrict__ x, const cplx<float> *__restrict__ d, int64_t n,
                              const double *__restrict__ nrm2v, const int *__restrict__ active)
{
    const int r = blockIdx.y;
    if (!active[r]) return;
    const double sc = sqrt(nrm2v[r]);
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const cplx<float> a = d[(int64_t)r * n + k];
        x[(int64_t)r * n + k] = x[(int64_t)r * n + k] + sc * cplx<double>{(double)a.re, (double)a.im};
    }
}

// FP32 copy of rows 1 and 2 of every link ([dir][e][site], e < 6): the FP32
// stencils rebuild row 3, so rows 3 of the copy are neither written nor read.
__global__ void k_links_to_f32(const cplx<double> *__restrict__ u, cplx<float> *__restrict__ v, int64_t vol)
{
    const int64_t n6 = 6 * vol;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < 3 * n6; j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = j + (j / n6) * 3 * vol; // dir * 9V + e * V + site
        v[k] = {(float)u[k].re, (float)u[k].im};
    }
}

// dst_r = src_r (n complex numbers per slot)
__global__ void k_copy_slots(slots<double> src, slots<double> dst, int64_t n)
{
    const int r = blockIdx.y;
    const cplx<double> *a = src.p[r];
    cplx<double> *d = dst.p[r];
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) d[k] = a[k];
}

// ---------------------------------------------------------------------------
// workspace
// ---------------------------------------------------------------------------
struct workspace {
    int64_t vol = 0;
    int m = 0;
    int64_t nsb = 0;                   // site blocks of the stencil launches
    int ext[3] = {0, 0, 0};            // lattice extents the buffers were sized for
    int rcap = 0;                      // slots the field buffers hold
    void *basis = nullptr;             // (m+1) x rcap x 3V complex double (or float)
    cplx<double> *y = nullptr, *x = nullptr, *t = nullptr, *res = nullptr;
    cplx<float> *y32 = nullptr, *x32 = nullptr, *t32 = nullptr, *res32 = nullptr;
    cplx<float> *links32 = nullptr;    // FP32 copy of the links for the inner solves
    cplx<double> *beta = nullptr, *cpar = nullptr, *hdev = nullptr, *ydev = nullptr, *cpart = nullptr;
    cplx<double> *negone = nullptr;    // -1 per slot: k_stencil with it as gco gives beta K in
    cplx<double> *cheb = nullptr;      // Chebyshev coefficients [iteration][slot][2]
    double *rpart = nullptr, *spart = nullptr, *nrm = nullptr;
    int *active = nullptr, *nused = nullptr;
    cplx<double> *fptr[NSLOTS] = {};   // heavy field of each slot in the current CN step
    cplx<double> *zbasis = nullptr;    // FGMRES: preconditioned directions, m x rcap x 3V
    int zm = 0;
    cplx<double> *hbuf = nullptr;      // pinned host
    double *nbuf = nullptr;            // pinned host (NRM doubles)

    // nrm layout, NSLOTS each: working norms, ||y||^2, saved refinement
    // ||r||^2, GMRES restart norms
    static constexpr int NRM_Y = NSLOTS, NRM_SAVED = 2 * NSLOTS, NRM_RESTART = 3 * NSLOTS, NRM = 4 * NSLOTS;

    void release()
    {
        void *dev[] = {basis, y, x, t, res, y32, x32, t32, res32, links32, beta, cpar, hdev, ydev, cpart,
                       rpart, spart, nrm, active, nused, cheb, negone};
        for (void *q : dev)
            if (q) (void)hipFree(q);
        if (hbuf) (void)hipHostFree(hbuf);
        if (nbuf) (void)hipHostFree(nbuf);
        if (zbasis) (void)hipFree(zbasis);
        *this = workspace();
    }

    // Buffers for lattice volume v, GMRES restart mm and at least r slots
    // (r = 0: the per-slot scalars only; field buffers are kept).
    int reserve(int64_t v, int mm, int r)
    {
        r = std::max(r, rcap);
        const geom &lg = lat().g;
        // the partial-sum buffer depends on the extents, not only on V
        const bool same = lg.n0 == ext[0] && lg.n1 == ext[1] && lg.n2 == ext[2];
        if (v == vol && mm == m && r == rcap && same) return NRQ_OK;
        release();
        const int rc = std::max(r, 1);
        const size_t f64 = (size_t)3 * (size_t)v * sizeof(cplx<double>) * rc;
        const size_t f32 = f64 / 2;
        int64_t sb = 0;
        for (int md = 0; md <= 4; ++md) sb = std::max(sb, make_lf_map(lg, md >= 3 ? SBS_WAVE : SBS, 8, md, 1).nsb);
        auto dm = [](void **p, size_t b) { return hipMalloc(p, b) == hipSuccess; };
        bool ok = dm(&basis, f64 * (size_t)(mm + 1)) && dm((void **)&y, f64) && dm((void **)&x, f64) &&
                  dm((void **)&t, f64) && dm((void **)&res, f64) && dm((void **)&y32, f32) &&
                  dm((void **)&x32, f32) && dm((void **)&t32, f32) && dm((void **)&res32, f32) &&
                  dm((void **)&links32, (size_t)27 * (size_t)v * sizeof(cplx<float>)) &&
                  dm((void **)&beta, sizeof(cplx<double>) * NSLOTS) && dm((void **)&cpar, sizeof(cplx<double>) * 2 * NSLOTS) &&
                  dm((void **)&negone, sizeof(cplx<double>) * NSLOTS) &&
                  dm((void **)&cheb, sizeof(cplx<double>) * 2 * NSLOTS * CHEB_MAX) &&
                  dm((void **)&hdev, sizeof(cplx<double>) * NSLOTS * (mm + 1)) &&
                  dm((void **)&ydev, sizeof(cplx<double>) * NSLOTS * (mm + 1)) &&
                  dm((void **)&cpart, sizeof(cplx<double>) * NSLOTS * (mm + 1) * VBLOCKS) &&
                  dm((void **)&rpart, sizeof(double) * std::max(KRYLOV_REDUCE_DOUBLES, (size_t)NSLOTS * VBLOCKS * 8)) &&
                  dm((void **)&spart, sizeof(double) * NSLOTS * (size_t)sb) && dm((void **)&nrm, sizeof(double) * NRM) &&
                  dm((void **)&active, sizeof(int) * NSLOTS) && dm((void **)&nused, sizeof(int) * NSLOTS) &&
                  hipHostMalloc((void **)&hbuf, sizeof(cplx<double>) * NSLOTS * (mm + 1), hipHostMallocDefault) == hipSuccess &&
                  hipHostMalloc((void **)&nbuf, sizeof(double) * NRM, hipHostMallocDefault) == hipSuccess;
        if (!ok) {
            (void)hipGetLastError();
            release();
            return fail(NRQ_ERR_ALLOC, "CN: workspace allocation failed (V=%lld, m=%d, %d slots)", (long long)v, mm, rc);
        }
        {
            std::vector<cplx<double>> m1(NSLOTS, cplx<double>{-1.0, 0.0});
            if (hipMemcpy(negone, m1.data(), sizeof(cplx<double>) * NSLOTS, hipMemcpyHostToDevice) != hipSuccess) {
                (void)hipGetLastError();
                release();
                return fail(NRQ_ERR_HIP, "CN: workspace initialisation failed");
            }
        }
        vol = v;
        m = mm;
        rcap = rc;
        nsb = sb;
        ext[0] = lg.n0;
        ext[1] = lg.n1;
        ext[2] = lg.n2;
        return NRQ_OK;
    }
};
static workspace g_ws;

int krylov_reserve(int64_t vol, int m, int slots) { return g_ws.reserve(vol, m, slots); }
const cplx<double> *krylov_last_rhs() { return g_ws.y; }

static int reserve_zbasis()
{
    if (g_ws.zbasis && g_ws.zm == g_ws.m) return NRQ_OK;
    if (g_ws.zbasis) (void)hipFree(g_ws.zbasis);
    g_ws.zbasis = nullptr;
    const size_t b = (size_t)g_ws.m * g_ws.rcap * 3 * (size_t)g_ws.vol * sizeof(cplx<double>);
    if (hipMalloc((void **)&g_ws.zbasis, b) != hipSuccess) {
        (void)hipGetLastError();
        return fail(NRQ_ERR_ALLOC, "CN: FGMRES basis allocation failed");
    }
    g_ws.zm = g_ws.m;
    return NRQ_OK;
}
void krylov_release() { g_ws.release(); }
double *krylov_reduce_scratch() { return g_ws.rpart; }
double *krylov_host_scratch() { return g_ws.nbuf; }

static int upload_ints(int *dev, const std::vector<int> &v)
{
    NRQ_HIP(hipMemcpyAsync(dev, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice, stream()));
    return NRQ_OK;
}
static int upload_active(const std::vector<int> &act) { return upload_ints(g_ws.active, act); }

// Launch shape of the stencil kernels: slots per thread (NC, sharing each
// link load) for FP64 and for FP32 arithmetic, and the site-block mapping of
// make_lf_map.  NRQ_CN_SHAPE="nc64,nc32,mode,xcd" in the environment sets
// the initial value, nrq_cn_set_shape changes it.  Every shape computes the
// same sums in the same order per slot.
struct cn_shape {
    // measured fastest at 128^3 with 12 fields: one work item per wavefront,
    // y-band z-sweep order, each XCD walking whole bands
    int nc64 = 2, nc32 = 4, mode = 4, xcd = 8;
};
static bool cn_nc_ok(int v) { return v == 1 || v == 2 || v == 4; }
static cn_shape read_shape()
{
    cn_shape sh;
    if (const char *e = getenv("NRQ_CN_SHAPE")) {
        int a = 0, b = 0, c = 0, d = 0;
        if (sscanf(e, "%d,%d,%d,%d", &a, &b, &c, &d) == 4) {
            if (cn_nc_ok(a)) sh.nc64 = a;
            if (cn_nc_ok(b)) sh.nc32 = b;
            if (c >= 0 && c <= 4) sh.mode = c;
            if (d == 1 || d == 8) sh.xcd = d;
        }
    }
    return sh;
}
static cn_shape &shape_mut()
{
    static cn_shape sh = read_shape();
    return sh;
}
static const cn_shape &shape() { return shape_mut(); }
template <typename T> static int nc_of() { return sizeof(T) == 8 ? shape().nc64 : shape().nc32; }

// stencil launch map for nf slots, nc per thread
static lf_map cn_map(int nf, int nc)
{
    return make_lf_map(lat().g, shape().mode >= 3 ? SBS_WAVE : SBS, shape().xcd, shape().mode, (nf + nc - 1) / nc);
}

// Runs the statements after `nc` with the compile-time chunk width NCV equal
// to nc (1, 2 or 4).
#define NC_DISPATCH(nc, ...)                                                                                             \
    do {                                                                                                                 \
        if ((nc) == 1) {                                                                                                 \
            constexpr int NCV = 1;                                                                                       \
            __VA_ARGS__;                                                                                                 \
        } else if ((nc) == 2) {                                                                                          \
            constexpr int NCV = 2;                                                                                       \
            __VA_ARGS__;                                                                                                 \
        } else {                                                                                                         \
            constexpr int NCV = 4;                                                                                       \
            __VA_ARGS__;                                                                                                 \
        }                                                                                                                \
    } while (0)

static unsigned grid_of(const lf_map &mp) { return lf_blocks(mp); }

// slot r -> base + r * 3V (workspace fields)
template <typename T> static slots<T> ws_slots(cplx<T> *base, int R)
{
    slots<T> t = {};
    for (int r = 0; r < R; ++r) t.p[r] = base + (int64_t)r * 3 * g_ws.vol;
    return t;
}
template <typename T> static slots<T> ws_slots(const cplx<T> *base, int R) { return ws_slots((cplx<T> *)base, R); }

// slot r -> the heavy field of the current CN step
static slots<double> psi_slots(int R)
{
    slots<double> t = {};
    for (int r = 0; r < R; ++r) t.p[r] = g_ws.fptr[r];
    return t;
}

// out[r] = sum of the nblk partials of slot r (r < nf; active may be null)
static int fold_slots(const double *partial, int64_t nblk, int nf, const int *active, double *out)
{
    hipLaunchKernelGGL(k_fold_slots, dim3(nf), dim3(FOLD_THREADS), 0, stream(), partial, nblk, active, out);
    NRQ_HIP(hipGetLastError());
    return NRQ_OK;
}

// g_ws.nrm to the pinned host buffer g_ws.nbuf; waits for the stream
static int nrm_to_host()
{
    NRQ_HIP(hipMemcpyAsync(g_ws.nbuf, g_ws.nrm, sizeof(double) * workspace::NRM, hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    return NRQ_OK;
}

// out_r = in_r + beta_r K in_r for the active slots
template <typename RL, typename T>
static void launch_stencil(const cplx<RL> *u, const slots<T> &in, const slots<T> &out, int nf,
                           const cplx<double> *gco = nullptr)
{
    const int nc = nc_of<T>();
    const lf_map mp = cn_map(nf, nc);
    NC_DISPATCH(nc, hipLaunchKernelGGL((k_stencil<NCV, RL, T>), dim3(grid_of(mp)), dim3(mp.bs), 0, stream(), u, in, out,
                                       g_ws.beta, g_ws.active, nf, lat().g, mp, gco));
}

// One Chebyshev step (k_cheb) for the active slots of R
template <typename RL, typename T>
static void launch_cheb(const cplx<RL> *links, const slots<T> &xk, const slots<T> &xnext, const slots<T> &b,
                        const cplx<double> *cf, int R, int mode, int fin, const slots<double> &fi, const slots<double> &fo)
{
    const int nc = nc_of<T>();
    const lf_map mp = cn_map(R, nc);
    NC_DISPATCH(nc, hipLaunchKernelGGL((k_cheb<NCV, RL, T>), dim3(grid_of(mp)), dim3(mp.bs), 0, stream(), links, xk, xnext,
                                       xnext, b, g_ws.beta, cf, g_ws.active, R, lat().g, mp, mode, fin, fi, fo));
}

// FP64 true residual r = y - A x of the R slots x (r stored as TR when
// WRITE); partial ||r||^2 to g_ws.spart.  Returns the partials per slot.
template <typename RL, typename TR, bool WRITE>
static int64_t launch_resid(const cplx<RL> *u, int R, const slots<double> &x, const slots<TR> &res)
{
    const slots<double> ty = ws_slots(g_ws.y, R);
    const int nc = shape().nc64;
    const lf_map mp = cn_map(R, nc);
    NC_DISPATCH(nc, hipLaunchKernelGGL((k_cn_resid<NCV, RL, TR, WRITE>), dim3(grid_of(mp)), dim3(mp.bs), 0, stream(), u, x, ty,
                                       res, g_ws.beta, R, lat().g, mp, g_ws.spart));
    return mp.nsb;
}

// Right-hand sides y and guesses x of a CN step (k_cn_rhs); partial ||y||^2
// to g_ws.spart.  Returns the partials per slot.
template <typename RL>
static int64_t launch_rhs(const cplx<RL> *u, int R, const slots<double> &psi, const slots<double> &y,
                          const slots<double> &x, int guess, const slots<float> &kout, int wk)
{
    const int nc = shape().nc64;
    const lf_map mp = cn_map(R, nc);
    NC_DISPATCH(nc, hipLaunchKernelGGL((k_cn_rhs<NCV, RL>), dim3(grid_of(mp)), dim3(mp.bs), 0, stream(), u, psi, y, x,
                                       g_ws.cpar, g_ws.cpar + NSLOTS, guess, R, lat().g, mp, g_ws.spart, kout, wk));
    return mp.nsb;
}

template <typename RL, typename T>
static int apply_t(const cplx<RL> *u, const cplx<T> *in, cplx<T> *out, int nf)
{
    launch_stencil<RL, T>(u, ws_slots(in, nf), ws_slots(out, nf), nf);
    NRQ_HIP(hipGetLastError());
    return NRQ_OK;
}

int krylov_norms2(const cplx<double> *x, int nf, double *out)
{
    std::vector<int> all(nf, 1);
    if (int e = upload_active(all)) return e;
    const int64_t n = 3 * g_ws.vol;
    hipLaunchKernelGGL(k_norm2<double>, dim3(VBLOCKS, nf), dim3(VTHREADS), 0, stream(), x, n, g_ws.active, g_ws.rpart);
    NRQ_HIP(hipGetLastError());
    if (int e = fold_slots(g_ws.rpart, VBLOCKS, nf, g_ws.active, g_ws.nrm)) return e;
    if (int e = nrm_to_host()) return e;
    memcpy(out, g_ws.nbuf, sizeof(double) * nf);
    return NRQ_OK;
}

// f1_r = beta_r K in_r, f2_r = beta_r K f1_r for the active slots of R
template <typename RL, typename T>
static int bk_features(const cplx<RL> *u, const slots<T> &in, const slots<T> &f1, const slots<T> &f2, int R)
{
    launch_stencil<RL, T>(u, in, f1, R, g_ws.negone);
    launch_stencil<RL, T>(u, f1, f2, R, g_ws.negone);
    NRQ_HIP(hipGetLastError());
    return NRQ_OK;
}

int krylov_rhs_features(const nrq_gauge *gf, int R, const cplx<double> **f1, const cplx<double> **f2)
{
    if (R <= 0 || R > g_ws.rcap) return fail(NRQ_ERR_ARG, "CN: %d features for a workspace of %d slots", R, g_ws.rcap);
    std::vector<int> all(R, 1);
    if (int e = upload_active(all)) return e;
    const slots<double> y = ws_slots(g_ws.y, R), a = ws_slots(g_ws.t, R), b = ws_slots(g_ws.res, R);
    const int e = gf->precision == 64 ? bk_features<double, double>((const cplx<double> *)gauge_links(gf), y, a, b, R)
                                      : bk_features<float, double>((const cplx<float> *)gauge_links(gf), y, a, b, R);
    if (e) return e;
    *f1 = g_ws.t;
    *f2 = g_ws.res;
    return NRQ_OK;
}

// column solves handed from FGMRES to plain GMRES since the library loaded
static int64_t g_fallbacks = 0;
// CN coefficients beta of the slots of the current step, on the host (the
// learned preconditioner scales its inputs and output per slot)
static cplx<double> g_beta_host[NSLOTS];

// ---------------------------------------------------------------------------
// GMRES(m) core: solves (1 + beta_r K) x_r = b_r for the slots with act[r],
// from the current x (x = 0 when zero_guess), until the T-precision true
// residual satisfies ||b - A x|| <= tol ||b|| or its[r] reaches maxit.
// rel_out[r] receives the last true relative residual.  beta must already be
// on the device.
// ---------------------------------------------------------------------------
// With net (FGMRES): z_j = M(v_j), w = A z_j, x += Z y.  A slot
// whose network output makes the Arnoldi step non-finite, whose cycle reduces
// its residual by less than FGMRES_STALL, or that has not converged after
// FGMRES_CYCLES cycles leaves the solve with x intact and fallen[r] = 1, for
// the caller to finish with plain GMRES.
constexpr double FGMRES_STALL = 0.1;
constexpr int FGMRES_CYCLES = 4;
template <typename T, typename RL>
static int gmres_core(const cplx<RL> *links, int R, const cplx<T> *b, cplx<T> *x, cplx<T> *basis, cplx<T> *t,
                      cplx<T> *res, bool zero_guess, const double *bnorm, std::vector<int> act, double tol, int m,
                      int64_t maxit, std::vector<int64_t> &its, double *rel_out, int64_t &cycles,
                      nrq_cnn *net = nullptr, cplx<T> *zbasis = nullptr, std::vector<int> *fallen = nullptr)
{
    typedef std::complex<double> zc;
    const int64_t n = 3 * g_ws.vol;
    const size_t fbytes = sizeof(cplx<T>) * (size_t)n;
    std::vector<double> rn(R, 0.0);
    if (int e = upload_active(act)) return e;
    const cplx<T> *r0 = res; // residual the next cycle starts from
    if (zero_guess) {
        NRQ_HIP(hipMemsetAsync(x, 0, fbytes * R, stream()));
        r0 = b;
        for (int r = 0; r < R; ++r) rn[r] = bnorm[r];
    } else {
        if (int e = apply_t<RL, T>(links, x, t, R)) return e;
        hipLaunchKernelGGL(k_residual<T>, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), b, t, res, n, g_ws.active, g_ws.rpart);
        if (int e = fold_slots(g_ws.rpart, VBLOCKS, R, g_ws.active, g_ws.nrm)) return e;
        if (int e = nrm_to_host()) return e;
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
    std::vector<int> jused(R, 0), cycle(R), ncycles(R, 0);
    std::vector<double> rn_start(R, 0.0);
    double *restart_n2 = g_ws.nrm + workspace::NRM_RESTART;
    while (true) {
        bool any = false;
        for (int r = 0; r < R; ++r) any = any || act[r];
        if (!any) break;
        ++cycles;
        rn_start = rn;
        std::vector<double> rn2(R);
        for (int r = 0; r < R; ++r) {
            rn2[r] = act[r] ? rn[r] * rn[r] : 0.0;
            std::fill(G[r].begin(), G[r].end(), zc(0));
            G[r][0] = rn[r];
            jused[r] = 0;
        }
        NRQ_HIP(hipMemcpyAsync(restart_n2, rn2.data(), sizeof(double) * R, hipMemcpyHostToDevice, stream()));
        if (int e = upload_active(act)) return e;
        hipLaunchKernelGGL(k_scale<T>, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), r0, basis, n, (const double *)restart_n2,
                           g_ws.active);
        cycle = act;
        for (int j = 0; j < m; ++j) {
            bool busy = false;
            for (int r = 0; r < R; ++r) busy = busy || cycle[r];
            if (!busy) break;
            if (int e = upload_active(cycle)) return e;
            cplx<T> *vj = basis + (size_t)j * R * n;
            cplx<T> *w = basis + (size_t)(j + 1) * R * n;
            if (net) {
                // z_j = M(v_j) from v_j and its features beta K v_j (in t) and
                // (beta K)^2 v_j (in res; the next cycle's residual is
                // recomputed after this loop)
                cplx<T> *zj = zbasis + (size_t)j * R * n;
                if (int e = bk_features<RL, T>(links, ws_slots(vj, R), ws_slots(t, R), ws_slots(res, R), R)) return e;
                const cplx<T> *vp[NSLOTS], *f1p[NSLOTS], *f2p[NSLOTS];
                cplx<T> *zp[NSLOTS];
                for (int r = 0; r < R; ++r) {
                    vp[r] = vj + (size_t)r * n;
                    f1p[r] = t + (size_t)r * n;
                    f2p[r] = res + (size_t)r * n;
                    zp[r] = zj + (size_t)r * n;
                }
                if (int e = cnn_precondition<T>(net, vp, f1p, f2p, zp, R, cycle.data(), g_beta_host)) return e;
                if (int e = apply_t<RL, T>(links, zj, w, R)) return e;
            } else if (int e = apply_t<RL, T>(links, vj, w, R)) {
                return e;
            }
            const int nv = j + 1;
            hipLaunchKernelGGL(k_mdot<T>, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), (const cplx<T> *)basis, nv,
                               (const cplx<T> *)w, n, R, g_ws.active, g_ws.cpart);
            hipLaunchKernelGGL(k_fold_c, dim3((nv * R + 3) / 4), dim3(256), 0, stream(), g_ws.cpart, VBLOCKS, nv, R, g_ws.active,
                               g_ws.hdev);
            hipLaunchKernelGGL(k_cgs_update<T>, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), (const cplx<T> *)basis, nv, w, n, R,
                               g_ws.active, g_ws.hdev, g_ws.rpart);
            if (int e = fold_slots(g_ws.rpart, VBLOCKS, R, g_ws.active, g_ws.nrm)) return e;
            hipLaunchKernelGGL(k_scale<T>, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), (const cplx<T> *)w, w, n,
                               (const double *)g_ws.nrm, g_ws.active);
            NRQ_HIP(hipGetLastError());
            NRQ_HIP(hipMemcpyAsync(g_ws.hbuf, g_ws.hdev, sizeof(cplx<double>) * nv * R, hipMemcpyDeviceToHost, stream()));
            if (int e = nrm_to_host()) return e;
            for (int r = 0; r < R; ++r) {
                if (!cycle[r]) continue;
                if (net) {
                    // the network's direction broke the Arnoldi step: end the
                    // slot's cycle before column j (x keeps columns < j)
                    bool fin = std::isfinite(g_ws.nbuf[r]);
                    for (int i = 0; i < nv && fin; ++i)
                        fin = std::isfinite(g_ws.hbuf[r * nv + i].re) && std::isfinite(g_ws.hbuf[r * nv + i].im);
                    if (!fin) {
                        cycle[r] = 0;
                        (*fallen)[r] = 1;
                        continue;
                    }
                }
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
            if (net) {
                bool fin = true;
                for (int i = 0; i < jj && fin; ++i) fin = std::isfinite(Y[r][i].real()) && std::isfinite(Y[r][i].imag());
                if (!fin) {
                    nu[r] = 0; // keep x
                    (*fallen)[r] = 1;
                }
            }
            for (int i = 0; i < nu[r]; ++i) g_ws.hbuf[r * (m + 1) + i] = {Y[r][i].real(), Y[r][i].imag()};
        }
        if (int e = upload_active(act)) return e;
        NRQ_HIP(hipMemcpyAsync(g_ws.ydev, g_ws.hbuf, sizeof(cplx<double>) * R * (m + 1), hipMemcpyHostToDevice, stream()));
        if (int e = upload_ints(g_ws.nused, nu)) return e;
        hipLaunchKernelGGL(k_xupdate<T>, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), x,
                           (const cplx<T> *)(net ? zbasis : basis), R, n, g_ws.ydev, m + 1, g_ws.nused, g_ws.active);
        // true residual in precision T
        if (int e = apply_t<RL, T>(links, x, t, R)) return e;
        hipLaunchKernelGGL(k_residual<T>, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), b, (const cplx<T> *)t, res, n,
                           g_ws.active, g_ws.rpart);
        r0 = res;
        if (int e = fold_slots(g_ws.rpart, VBLOCKS, R, g_ws.active, g_ws.nrm)) return e;
        if (int e = nrm_to_host()) return e;
        for (int r = 0; r < R; ++r) {
            if (!act[r]) continue;
            rn[r] = std::sqrt(std::max(g_ws.nbuf[r], 0.0));
            rel_out[r] = bnorm[r] > 0 ? rn[r] / bnorm[r] : 0.0;
            if (rn[r] <= tol * bnorm[r] || its[r] >= maxit) {
                act[r] = 0;
                continue;
            }
            if (net && ((*fallen)[r] || !(rn[r] <= FGMRES_STALL * rn_start[r]) || ++ncycles[r] >= FGMRES_CYCLES)) {
                act[r] = 0;
                (*fallen)[r] = 1;
            }
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
typedef std::complex<double> zc;

static double cheb_kmax()
{
    const geom &g = lat().g;
    return 2.0 / (g.a[1] * g.a[1]) + 2.0 / (g.a[2] * g.a[2]) + 2.0 / (g.a[3] * g.a[3]);
}

// 1 / |T_k(sigma)|
static double cheb_bound(zc sg, int k)
{
    zc t0 = 1.0, t1 = sg;
    for (int j = 1; j < k; ++j) {
        const zc t2 = 2.0 * sg * t1 - t0;
        t0 = t1;
        t1 = t2;
    }
    return 1.0 / std::abs(t1);
}

// Smallest k < CHEB_MAX with 1/|T_k(sigma)| <= target (CHEB_MAX if none).
static int cheb_steps(zc sg, double target)
{
    for (int k = 1; k < CHEB_MAX; ++k)
        if (cheb_bound(sg, k) <= target) return k;
    return CHEB_MAX;
}

// Bound on ||y - A x0|| / ||y|| for the initial guess.  With the first-order
// guess x0 = psi + (alpha - beta) K psi the residual is
// -beta (alpha - beta) K^2 psi, and ||y|| >= min_{0<=l<=kmax} |1 + alpha l| ||psi||.
static double guess_bound(zc a, zc b, bool guess, double kmax)
{
    if (!guess) return 1.0;
    const double lam = a.real() < 0 ? std::min(kmax, -a.real() / std::norm(a)) : 0.0;
    const double mn = std::abs(1.0 + a * lam);
    if (!(mn > 0)) return 1.0;
    return std::min(1.0, 1.05 * std::abs(b) * std::abs(a - b) * kmax * kmax / mn);
}

// Coefficient table for k steps; tab[j] = {1/c_r, rho_r of step j}.
static int cheb_table(int R, const std::vector<zc> &beta, int k, double kmax)
{
    std::vector<cplx<double>> tab((size_t)2 * NSLOTS * k, cplx<double>{0.0, 0.0});
    for (int r = 0; r < R; ++r) {
        const zc c = 1.0 + beta[r] * (kmax / 2), d = beta[r] * (kmax / 2);
        const zc sg = c / d, ic = 1.0 / c;
        zc rho = 1.0;
        for (int j = 0; j < k; ++j) {
            if (j == 1) rho = 1.0 / (1.0 - 1.0 / (2.0 * sg * sg));
            else if (j > 1) rho = 1.0 / (1.0 - rho / (4.0 * sg * sg));
            tab[(size_t)2 * NSLOTS * j + 2 * r] = {ic.real(), ic.imag()};
            tab[(size_t)2 * NSLOTS * j + 2 * r + 1] = {rho.real(), rho.imag()};
        }
    }
    NRQ_HIP(hipMemcpyAsync(g_ws.cheb, tab.data(), sizeof(cplx<double>) * tab.size(), hipMemcpyHostToDevice, stream()));
    return NRQ_OK;
}

// k Chebyshev steps on A x = b for the active slots.  zero_start: x starts at
// 0 and steps 0 and 1 are folded into the first launch (k >= 2); otherwise
// from the iterate in `x`.  x and xalt alternate; *out receives the buffer
// holding x_k.  With fin_in/fin_out the last step instead stores
// fin_out = fin_in + x_k in FP64.
template <typename RL, typename T>
static int cheb_run(const cplx<RL> *links, int R, const cplx<T> *b, cplx<T> *x, cplx<T> *xalt, bool zero_start,
                    const std::vector<zc> &beta, int k, double kmax, cplx<T> **out, const slots<double> *fin_in = nullptr,
                    const slots<double> *fin_out = nullptr)
{
    if (zero_start && k < 2) k = 2;
    if (k > CHEB_MAX) k = CHEB_MAX;
    if (int e = cheb_table(R, beta, k, kmax)) return e;
    const slots<T> tb = ws_slots(b, R);
    const slots<double> none = {};
    cplx<T> *cur = x, *other = xalt;
    for (int j = zero_start ? 1 : 0; j < k; ++j) {
        const int mode = zero_start ? (j == 1 ? 2 : (j == 2 ? 3 : 0)) : (j == 0 ? 1 : 0);
        // mode 2 reads b and writes cur; the other modes read cur and write other
        const slots<T> txk = mode == 2 ? tb : ws_slots(cur, R);
        const slots<T> tnext = ws_slots(mode == 2 ? cur : other, R);
        const int fin = (fin_in && j == k - 1) ? 1 : 0;
        const slots<double> fi = fin ? *fin_in : none, fo = fin ? *fin_out : none;
        const cplx<double> *cf = g_ws.cheb + (size_t)2 * NSLOTS * j;
        launch_cheb<RL, T>(links, txk, tnext, tb, cf, R, mode, fin, fi, fo);
        if (mode != 2) std::swap(cur, other);
    }
    NRQ_HIP(hipGetLastError());
    *out = cur;
    return NRQ_OK;
}

// FP64 true residual of the slot fields x with its norms into g_ws.nrm[0, R);
// res32 (may be null) receives the residual in FP32.
template <typename RL>
static int cn_resid(const cplx<RL> *u, int R, const slots<double> &x, cplx<float> *res32)
{
    int64_t nsb;
    if (res32)
        nsb = launch_resid<RL, float, true>(u, R, x, ws_slots(res32, R));
    else
        nsb = launch_resid<RL, float, false>(u, R, x, slots<float>{});
    NRQ_HIP(hipGetLastError());
    return fold_slots(g_ws.spart, nsb, R, nullptr, g_ws.nrm);
}

// Per-slot convergence check of the residual norms in g_ws.nbuf; returns
// true if a slot is still active.  need[r] receives the residual reduction
// still needed.
static bool check_slots(int R, const std::vector<double> &bn, double tol, std::vector<int> &act, std::vector<double> &rel,
                        std::vector<double> &need)
{
    bool any = false;
    for (int r = 0; r < R; ++r) {
        if (!act[r]) continue;
        const double rr = std::sqrt(std::max(g_ws.nbuf[r], 0.0));
        rel[r] = bn[r] > 0 ? rr / bn[r] : 0.0;
        if (rr <= tol * bn[r]) {
            act[r] = 0;
            continue;
        }
        any = true;
        need[r] = 0.3 * tol * bn[r] / rr;
    }
    return any;
}

// Chebyshev solve from the guess in g_ws.x.  On return *in_psi tells whether
// the iterate is already in the heavy field (else in *xfinal).  Slots still
// active on return go on with GMRES.
template <typename RL>
static int cheb_solve(const cplx<RL> *u, int R, const std::vector<zc> &alpha, const std::vector<zc> &beta,
                      std::vector<double> &bn, const krylov_params &p, std::vector<int> &act, std::vector<double> &rel,
                      std::vector<int64_t> &its, int64_t &cycles, cplx<double> **xfinal, bool *in_psi)
{
    const double kmax = cheb_kmax();
    std::vector<zc> sig(R);
    std::vector<double> b0(R);
    for (int r = 0; r < R; ++r) {
        const zc d = beta[r] * (kmax / 2);
        sig[r] = std::abs(d) > 0 ? (1.0 + d) / d : zc(1e30, 0.0);
        b0[r] = guess_bound(alpha[r], beta[r], p.extrapolate, kmax);
    }
    std::vector<double> need(R, 1.0);
    auto steps_for = [&](double floor_) {
        int k = 1;
        for (int r = 0; r < R; ++r)
            if (act[r]) k = std::max(k, cheb_steps(sig[r], std::max(need[r], floor_)));
        return k;
    };
    auto bn_from_host = [&]() {
        for (int r = 0; r < R; ++r) bn[r] = std::sqrt(std::max(g_ws.nbuf[workspace::NRM_Y + r], 0.0));
    };
    *in_psi = false;
    if (p.kind == KRYLOV_CHEB_FP64) {
        // the first round takes enough steps for the a priori guess bound
        for (int r = 0; r < R; ++r) need[r] = 0.3 * p.tol / b0[r];
        cplx<double> *x = g_ws.x;
        for (int round = 0; round < 8; ++round) {
            const int k = steps_for(0.0);
            ++cycles;
            cplx<double> *xk = nullptr;
            if (int e = cheb_run<RL, double>(u, R, g_ws.y, x, x == g_ws.x ? g_ws.t : g_ws.x, false, beta, k, kmax, &xk)) return e;
            x = xk;
            for (int r = 0; r < R; ++r) its[r] += act[r] ? k : 0;
            if (int e = cn_resid<RL>(u, R, ws_slots(x, R), nullptr)) return e;
            if (int e = nrm_to_host()) return e;
            if (round == 0) bn_from_host();
            if (!check_slots(R, bn, p.tol, act, rel, need)) break;
            if (int e = upload_active(act)) return e;
        }
        *xfinal = x;
        return NRQ_OK;
    }
    // mixed: FP64 residual -> FP32 Chebyshev correction d from zero whose
    // last step stores x + d into the heavy field (from the guess the first
    // time, in place afterwards)
    const cplx<float> *l32;
    if (sizeof(RL) == 8) {
        hipLaunchKernelGGL(k_links_to_f32, dim3(4096), dim3(256), 0, stream(), (const cplx<double> *)u, g_ws.links32, g_ws.vol);
        NRQ_HIP(hipGetLastError());
        l32 = g_ws.links32;
    } else {
        l32 = (const cplx<float> *)u;
    }
    const slots<double> tpsi = psi_slots(R), tx = ws_slots(g_ws.x, R);
    if (p.extrapolate) {
        // the guess's residual -beta g K^2 psi from K psi (FP32, left in t32
        // by the right-hand-side pass): one FP32 stencil pass instead of an
        // FP64 residual pass; the FP64 true residual after the correction
        // still decides
        launch_stencil<float, float>(l32, ws_slots((const cplx<float> *)g_ws.t32, R), ws_slots(g_ws.y32, R), R,
                                     g_ws.cpar + NSLOTS);
        NRQ_HIP(hipGetLastError());
    } else if (int e = cn_resid<RL>(u, R, tx, g_ws.y32)) {
        return e;
    }
    auto correct = [&](int k) -> int {
        cplx<float> *d = nullptr;
        const slots<double> &from = *in_psi ? tpsi : tx;
        if (int e = cheb_run<float, float>(l32, R, g_ws.y32, g_ws.x32, g_ws.t32, true, beta, k, kmax, &d, &from, &tpsi))
            return e;
        *in_psi = true;
        return cn_resid<RL>(u, R, tpsi, g_ws.y32);
    };
    // corrections planned from the a priori bound, without synchronisation
    std::vector<double> bnd = b0;
    for (int plan = 0; plan < 3; ++plan) {
        bool more = false;
        for (int r = 0; r < R; ++r) {
            need[r] = bnd[r] > 0.3 * p.tol ? 0.3 * p.tol / bnd[r] : 1.0;
            more = more || bnd[r] > 0.3 * p.tol;
        }
        if (!more) break;
        const int k = std::max(2, steps_for(3e-7));
        ++cycles;
        if (int e = correct(k)) return e;
        for (int r = 0; r < R; ++r) {
            its[r] += k;
            bnd[r] *= cheb_bound(sig[r], k) + FP32_CORR_ERR;
        }
    }
    if (int e = nrm_to_host()) return e;
    bn_from_host();
    if (!*in_psi) {
        // no correction was needed: the guess goes to the heavy field
        hipLaunchKernelGGL(k_copy_slots, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), tx, tpsi, 3 * g_ws.vol);
        NRQ_HIP(hipGetLastError());
        *in_psi = true;
    }
    // measured corrections (in place) until converged
    for (int round = 0; round < 8; ++round) {
        if (!check_slots(R, bn, p.tol, act, rel, need)) break;
        if (int e = upload_active(act)) return e;
        const int k = std::max(2, steps_for(3e-7));
        ++cycles;
        if (int e = correct(k)) return e;
        for (int r = 0; r < R; ++r) its[r] += act[r] ? k : 0;
        if (int e = nrm_to_host()) return e;
    }
    return NRQ_OK;
}

// FP64 iterative refinement around FP32 GMRES corrections, from the guess in
// g_ws.x; with net, the corrections are FP32 FGMRES with the network (slots
// the network does not help finish their correction with plain FP32 GMRES).
template <typename RL>
static int gmres_mixed(const cplx<RL> *u, int R, std::vector<double> &bn, const krylov_params &p, std::vector<int> &act,
                       std::vector<double> &rel, std::vector<int64_t> &its, int64_t &cycles, nrq_cnn *net = nullptr)
{
    const int64_t n = 3 * g_ws.vol;
    const cplx<float> *l32;
    if (sizeof(RL) == 8) {
        hipLaunchKernelGGL(k_links_to_f32, dim3(4096), dim3(256), 0, stream(), (const cplx<double> *)u, g_ws.links32, g_ws.vol);
        NRQ_HIP(hipGetLastError());
        l32 = g_ws.links32;
    } else {
        l32 = (const cplx<float> *)u;
    }
    std::vector<double> one(R, 1.0);
    const slots<double> tx = ws_slots(g_ws.x, R), tres = ws_slots(g_ws.res, R);
    for (int outer = 0; outer < 40; ++outer) {
        // FP64 true residual r = y - A x
        const int64_t nsb = launch_resid<RL, double, true>(u, R, tx, tres);
        NRQ_HIP(hipGetLastError());
        if (int e = fold_slots(g_ws.spart, nsb, R, nullptr, g_ws.nrm)) return e;
        if (int e = nrm_to_host()) return e;
        bool any = false;
        for (int r = 0; r < R; ++r) {
            if (!act[r]) continue;
            const double rr = std::sqrt(std::max(g_ws.nbuf[r], 0.0));
            rel[r] = bn[r] > 0 ? rr / bn[r] : 0.0;
            if (rr <= p.tol * bn[r] || its[r] >= p.maxit) act[r] = 0;
            any = any || act[r];
        }
        if (!any) break;
        ++cycles;
        // ||r||^2 of this refinement step at nrm[64..): the inner solve reuses nrm[0..R)
        double *rnorm2 = g_ws.nrm + workspace::NRM_SAVED;
        NRQ_HIP(hipMemcpyAsync(rnorm2, g_ws.nrm, sizeof(double) * R, hipMemcpyDeviceToDevice, stream()));
        // normalised FP32 correction problem A d = r / ||r||
        if (int e = upload_active(act)) return e;
        hipLaunchKernelGGL(k_demote, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), (const cplx<double> *)g_ws.res, g_ws.y32, n,
                           (const double *)rnorm2, g_ws.active);
        std::vector<double> inner_rel(R, 0.0);
        int64_t inner_cycles = 0;
        const double itol = std::max(p.inner_tol, std::max(p.tol / 1e-3, 2e-7));
        std::vector<int> fallen(R, 0);
        if (int e = gmres_core<float, float>(l32, R, g_ws.y32, g_ws.x32, (cplx<float> *)g_ws.basis, g_ws.t32, g_ws.res32, true,
                                             one.data(), act, itol, p.restart, p.maxit, its, inner_rel.data(), inner_cycles,
                                             net, (cplx<float> *)g_ws.zbasis, net ? &fallen : nullptr))
            return e;
        int nfall = 0;
        for (int f : fallen) nfall += f;
        if (nfall) {
            g_fallbacks += nfall;
            if (int e = gmres_core<float, float>(l32, R, g_ws.y32, g_ws.x32, (cplx<float> *)g_ws.basis, g_ws.t32, g_ws.res32,
                                                 false, one.data(), fallen, itol, p.restart, p.maxit, its, inner_rel.data(),
                                                 inner_cycles))
                return e;
        }
        // gmres_core left the slot masks of the inner solve on the device;
        // the refinement updates every slot that is still active here
        if (int e = upload_active(act)) return e;
        hipLaunchKernelGGL(k_promote_add, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), g_ws.x, (const cplx<float> *)g_ws.x32, n,
                           (const double *)rnorm2, g_ws.active);
        NRQ_HIP(hipGetLastError());
    }
    return NRQ_OK;
}

template <typename RL>
static int cn_step_t(const cplx<RL> *u, int R, const std::vector<zc> &alpha, const std::vector<zc> &beta,
                     const krylov_params &p, nrq_solve_stats *st)
{
    const int64_t n = 3 * g_ws.vol;
    std::vector<double> bn(R, 0.0), rel(R, 0.0);
    std::vector<int64_t> its(R, 0);
    int64_t cycles = 0;
    std::vector<int> act(R, 1);
    const bool guess = p.extrapolate && p.kind != KRYLOV_GMRES_FP64;
    // right-hand sides and guesses; ||y||^2 to nrm[32..)
    const slots<double> tpsi = psi_slots(R), ty = ws_slots(g_ws.y, R), tx = ws_slots(g_ws.x, R);
    // mixed Chebyshev from the guess takes K psi (FP32) for its first correction
    const int wk = guess && p.kind == KRYLOV_CHEB_MIXED ? 1 : 0;
    const int64_t nsb = launch_rhs<RL>(u, R, tpsi, ty, tx, guess ? 1 : 0, ws_slots(g_ws.t32, R), wk);
    NRQ_HIP(hipGetLastError());
    if (int e = fold_slots(g_ws.spart, nsb, R, nullptr, g_ws.nrm + workspace::NRM_Y)) return e;
    cplx<double> *xsol = g_ws.x;
    bool in_psi = false;
    int status = NRQ_OK;
    if (p.kind == KRYLOV_CHEB_FP64 || p.kind == KRYLOV_CHEB_MIXED) {
        if (int e = cheb_solve<RL>(u, R, alpha, beta, bn, p, act, rel, its, cycles, &xsol, &in_psi)) return e;
        bool left = false;
        for (int r = 0; r < R; ++r) left = left || act[r];
        if (left) {
            // spectrum bound violated (e.g. links far from unitary): GMRES from here
            if (in_psi) {
                hipLaunchKernelGGL(k_copy_slots, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), tpsi, tx, n);
                NRQ_HIP(hipGetLastError());
                in_psi = false;
            } else if (xsol != g_ws.x) {
                NRQ_HIP(hipMemcpyAsync(g_ws.x, xsol, sizeof(cplx<double>) * (size_t)n * R, hipMemcpyDeviceToDevice, stream()));
            }
            xsol = g_ws.x;
            status = gmres_core<double, RL>(u, R, g_ws.y, g_ws.x, (cplx<double> *)g_ws.basis, g_ws.t, g_ws.res, false,
                                            bn.data(), act, p.tol, p.restart, p.maxit, its, rel.data(), cycles);
        }
    } else {
        if (int e = nrm_to_host()) return e;
        for (int r = 0; r < R; ++r) bn[r] = std::sqrt(std::max(g_ws.nbuf[workspace::NRM_Y + r], 0.0));
        if (p.kind == KRYLOV_GMRES_FP64)
            status = gmres_core<double, RL>(u, R, g_ws.y, g_ws.x, (cplx<double> *)g_ws.basis, g_ws.t, g_ws.res, true,
                                            bn.data(), act, p.tol, p.restart, p.maxit, its, rel.data(), cycles);
        else if (p.kind == KRYLOV_FGMRES_CNN) {
            nrq_cnn *net = cn_preconditioner();
            if (!net) return fail(NRQ_ERR_STATE, "CN: FGMRES needs a network (nrq_cn_set_preconditioner)");
            if (int e = reserve_zbasis()) return e;
            std::vector<int> fallen(R, 0);
            status = gmres_core<double, RL>(u, R, g_ws.y, g_ws.x, (cplx<double> *)g_ws.basis, g_ws.t, g_ws.res,
                                            !p.extrapolate, bn.data(), act, p.tol, p.restart, p.maxit, its, rel.data(),
                                            cycles, net, g_ws.zbasis, &fallen);
            int nfall = 0;
            for (int f : fallen) nfall += f;
            if (!status && nfall) {
                // the network did not help these slots: plain GMRES (the
                // operator's diagonal is constant, so Jacobi scaling leaves
                // GMRES unchanged) from the iterate reached so far
                g_fallbacks += nfall;
                status = gmres_core<double, RL>(u, R, g_ws.y, g_ws.x, (cplx<double> *)g_ws.basis, g_ws.t, g_ws.res,
                                                false, bn.data(), fallen, p.tol, p.restart, p.maxit, its, rel.data(),
                                                cycles);
            }
        } else if (p.kind == KRYLOV_FGMRES_CNN_MIXED) {
            nrq_cnn *net = cn_preconditioner();
            if (!net) return fail(NRQ_ERR_STATE, "CN: FGMRES needs a network (nrq_cn_set_preconditioner)");
            if (int e = reserve_zbasis()) return e;
            status = gmres_mixed<RL>(u, R, bn, p, act, rel, its, cycles, net);
        } else
            status = gmres_mixed<RL>(u, R, bn, p, act, rel, its, cycles);
    }
    if (status) return status;
    if (!in_psi) {
        // psi_f = x for every slot
        hipLaunchKernelGGL(k_copy_slots, dim3(VBLOCKS, R), dim3(VTHREADS), 0, stream(), ws_slots(xsol, R), tpsi, n);
        NRQ_HIP(hipGetLastError());
    }
    double worst = 0;
    int64_t total = 0;
    for (int r = 0; r < R; ++r) {
        worst = std::max(worst, rel[r]);
        total += its[r];
    }
    if (st) {
        st->iterations += total;
        st->restarts += cycles;
        st->max_true_relres = std::max(st->max_true_relres, worst);
    }
    if (worst > p.tol)
        return fail(NRQ_ERR_NOCONV, "CN: Krylov solver stopped (maxit=%lld) with true relative residual %.3e > %.1e",
                    (long long)p.maxit, worst, p.tol);
    return NRQ_OK;
}

int krylov_cn_step(const nrq_gauge *gf, cplx<double> *const *fields, int R, const cplx<double> *alpha,
                   const cplx<double> *beta, const krylov_params &p, nrq_solve_stats *st)
{
    if (R <= 0) return NRQ_OK;
    if (R > NSLOTS) return fail(NRQ_ERR_ARG, "CN: %d fields exceed the batch of %d", R, NSLOTS);
    if (g_ws.vol != lat().g.vol || g_ws.rcap < R) return fail(NRQ_ERR_STATE, "CN: workspace not reserved for this solve");
    std::vector<int> act(R, 1);
    std::vector<cplx<double>> par(2 * NSLOTS, cplx<double>{0.0, 0.0});
    std::vector<zc> a(R), b(R);
    for (int r = 0; r < R; ++r) {
        if (!fields[r]) return fail(NRQ_ERR_ARG, "CN: null field %d", r);
        g_ws.fptr[r] = fields[r];
        a[r] = zc(alpha[r].re, alpha[r].im);
        b[r] = zc(beta[r].re, beta[r].im);
        par[r] = alpha[r];
        g_beta_host[r] = beta[r];
        const zc g = a[r] - b[r];
        par[NSLOTS + r] = {g.real(), g.imag()};
    }
    if (int e = upload_active(act)) return e;
    NRQ_HIP(hipMemcpyAsync(g_ws.cpar, par.data(), sizeof(cplx<double>) * par.size(), hipMemcpyHostToDevice, stream()));
    NRQ_HIP(hipMemcpyAsync(g_ws.beta, beta, sizeof(cplx<double>) * R, hipMemcpyHostToDevice, stream()));
    if (gf->precision == 64) return cn_step_t<double>((const cplx<double> *)gauge_links(gf), R, a, b, p, st);
    return cn_step_t<float>((const cplx<float> *)gauge_links(gf), R, a, b, p, st);
}

} // namespace nrq

using namespace nrq;

extern "C" int nrq_cn_set_shape(int nc64, int nc32, int tiling, int xcd_remap)
{
    if (!cn_nc_ok(nc64) || !cn_nc_ok(nc32)) return fail(NRQ_ERR_ARG, "CN shape: slots per thread must be 1, 2 or 4");
    if (tiling < 0 || tiling > 4) return fail(NRQ_ERR_ARG, "CN shape: tiling %d out of range", tiling);
    if (xcd_remap != 1 && xcd_remap != 8) return fail(NRQ_ERR_ARG, "CN shape: XCD remap must be 1 or 8");
    cn_shape &sh = shape_mut();
    sh.nc64 = nc64;
    sh.nc32 = nc32;
    sh.mode = tiling;
    sh.xcd = xcd_remap;
    return NRQ_OK;
}

extern "C" int nrq_cn_get_shape(int *nc64, int *nc32, int *tiling, int *xcd_remap)
{
    if (!nc64 || !nc32 || !tiling || !xcd_remap) return fail(NRQ_ERR_ARG, "null argument");
    const cn_shape &sh = shape();
    *nc64 = sh.nc64;
    *nc32 = sh.nc32;
    *tiling = sh.mode;
    *xcd_remap = sh.xcd;
    return NRQ_OK;
}

extern "C" int nrq_cn_fallbacks(int64_t *count)
{
    if (!count) return fail(NRQ_ERR_ARG, "null argument");
    *count = g_fallbacks;
    return NRQ_OK;
}
