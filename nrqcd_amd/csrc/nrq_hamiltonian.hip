// nrq_hamiltonian.hip -- the Crank-Nicolson step with the full NRQCD
// Hamiltonian of BuildSystemMatrix (nrqcd.f90:1175-1387): the kinetic c1 term
// plus the c2 (magnetic), c3 (E derivative) and c4 (spin-orbit) terms, taken
// when any of c2..c4 is non-zero.
//
// S = 1 + sum_t a_t T_t with (dt the step width, pre = -i dt a0 / 2)
//   T_1 psi = sum_i [psi(x) - (U_i(x) psi(x+i) + U_i(x-i)^+ psi(x-i))/2] / a_i^2,
//            a_1 = pre c1 / M;
//   T_2 = C2CS(Bsum(x)),  Bsum = sum_i (F_jk - F_kj)/2 (nrqcd.f90:1669-1693; no
//            Pauli matrix, a reference quirk), a_2 = pre (-c2) / (2M);
//   T_3 = C2CS(Dsum(x)),  Dsum = sum_i (U_i(x) e_i(x+i) U_i(x)^+
//            - U_i(x-i)^+ e_i(x-i) U_i(x-i)) / (2 a_i), e = -F_0i,
//            a_3 = pre (-c3) / (8 M^2);
//   T_4 hops to x +- i with Kron(G, sigma_k) for the cyclic (eps = +1) and
//            anticyclic (eps = -1) (i, j, k):
//            to x + i: G = eps (e_j(x) U_i(x) + U_i(x) e_j(x+i)) / (2 a_i),
//            to x - i: G = -eps (e_j(x) W + W e_j(x-i)) / (2 a_i), W = U_i(x-i)^+,
//            a_4 = pre (-c4) i / (8 M^2).
// The field strength F_ij = Re[ArgLogU(P_ij)] / (a_i a_j) is the element-wise
// real part (gaugeconfiguration_su3.f90:1563-1581) with the spatial plaquette
// or, for F_0j, the temporal plaquette exp(i E_j^a T^a) (:1656-1679).  The
// antiquark uses (-M, conj c) (nrqcd.f90:1130-1161).
//
// Per step: rhs = S psi, then S^+ psi' = rhs (S^+ = 1 + sum_t conj(a_t) T_t^+)
// for every column of both propagators.  c4 mixes the spin rows, so a column
// is the pair of colour-vector fields (2j, 2j+1), 6V complex doubles,
// contiguous in the field storage.  The solver is FP64 GMRES(m) with
// classical Gram-Schmidt and zero initial guess (PETSc's defaults,
// nrqcd.f90:803-804) that stops on the FP64 true relative residual.
//
// The field strengths (kernel k_ham_fields: 3 temporal and 6 spatial
// logarithms per site) and the colour matrices Dsum and G (k_ham_blocks) are
// built once per step; k_ham_apply applies S or S^+ to several columns at
// once.
#include "nrq_internal.h"
#include "nrq_grouplog.h"
#include "nrq_map.h"
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <complex>
#include <vector>

namespace nrq {

const void *gauge_links(const nrq_gauge *h);

constexpr int HAM_THREADS = 256;
constexpr int HAM_FLD = 36;   // doubles per site: F0_j (3 x 9), Bsum (9)
constexpr int HAM_BLK = 13;   // colour matrices per site: Dsum, G[i][sign][p]
constexpr int HAM_MAXCOL = 12; // columns per apply launch (6 per propagator)
constexpr size_t HAM_BATCH_BYTES = (size_t)16 << 30; // Krylov storage of the columns solved together
constexpr int HAM_DOT_BLOCKS = 512;

template <typename R> __device__ __forceinline__ mat3<double> ham_link(const cplx<R> *u, int64_t vol, int dir, int64_t s)
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

__device__ __forceinline__ mat3<double> madd(const mat3<double> &a, const mat3<double> &b, double sb)
{
    mat3<double> r;
#pragma unroll
    for (int e = 0; e < 9; ++e) r.m[e] = {a.m[e].re + sb * b.m[e].re, a.m[e].im + sb * b.m[e].im};
    return r;
}

// F0_j(x) and Bsum(x) into fld[(q * 9 + e) V + s], q = 0..2 for F0_j, 3 for Bsum.
template <typename R>
__global__ __launch_bounds__(HAM_THREADS) void k_ham_fields(const cplx<R> *__restrict__ u, const R *__restrict__ ef, geom g,
                                                           double *__restrict__ fld)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.vol) return;
    const int64_t V = g.vol;
    for (int j = 0; j < 3; ++j) {
        double al[8];
        for (int a = 0; a < 8; ++a) al[a] = (double)ef[((int64_t)j * 8 + a) * V + s];
        const mat3<double> H = group_log_matrix(exp_i_hermitian(algebra_matrix(al)));
        const double sc = 1.0 / (g.a[0] * g.a[j + 1]);
        for (int e = 0; e < 9; ++e) fld[((int64_t)j * 9 + e) * V + s] = H.m[e].re * sc;
    }
    const lsite p = make_site(g, s);
    double b[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3, k = (i + 2) % 3;
        for (int o = 0; o < 2; ++o) {
            // P_mn(x) = U_m(x) U_n(x+m) U_m(x+n)^+ U_n(x)^+ for (m, n) = (j, k), then (k, j)
            const int m = o == 0 ? j : k, n = o == 0 ? k : j;
            const mat3<double> P = mul_nd(mul_nd(mul(ham_link(u, V, m, s), ham_link(u, V, n, site_up(g, p, m).s)),
                                                 ham_link(u, V, m, site_up(g, p, n).s)),
                                          ham_link(u, V, n, s));
            const mat3<double> H = group_log_matrix(P);
            const double sc = (o == 0 ? 0.5 : -0.5) / (g.a[m + 1] * g.a[n + 1]);
            for (int e = 0; e < 9; ++e) b[e] += H.m[e].re * sc;
        }
    }
    for (int e = 0; e < 9; ++e) fld[((int64_t)27 + e) * V + s] = b[e];
}

__device__ __forceinline__ mat3<double> ham_e(const double *__restrict__ fld, int64_t V, int j, int64_t s)
{
    mat3<double> m;
#pragma unroll
    for (int e = 0; e < 9; ++e) m.m[e] = {-fld[((int64_t)j * 9 + e) * V + s], 0.0};
    return m;
}

// Dsum (matrix 0) and G[i][sign][p] (matrix 1 + (2 i + sign) 2 + p) into
// blk[(mat * 9 + e) V + s].
template <typename R>
__global__ __launch_bounds__(HAM_THREADS) void k_ham_blocks(const cplx<R> *__restrict__ u, const double *__restrict__ fld,
                                                           geom g, cplx<double> *__restrict__ blk)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.vol) return;
    const int64_t V = g.vol;
    const lsite p = make_site(g, s);
    mat3<double> dsum;
    set_zero(dsum);
    for (int i = 0; i < 3; ++i) {
        const int64_t sp = site_up(g, p, i).s, sm = site_dn(g, p, i).s;
        const double inv2a = 1.0 / (2.0 * g.a[i + 1]);
        const mat3<double> Ui = ham_link(u, V, i, s), Um = ham_link(u, V, i, sm);
        // c3: U e_i(x+i) U^+ - Um^+ e_i(x-i) Um
        const mat3<double> t1 = mul_nd(mul(Ui, ham_e(fld, V, i, sp)), Ui);
        const mat3<double> t2 = mul(mul_dn(Um, ham_e(fld, V, i, sm)), Um);
        dsum = madd(dsum, madd(t1, t2, -1.0), inv2a);
        // c4
        mat3<double> W;
        for (int e = 0; e < 9; ++e) W.m[e] = {Um.m[3 * (e % 3) + e / 3].re, -Um.m[3 * (e % 3) + e / 3].im};
        for (int q = 0; q < 2; ++q) {
            const int j = q == 0 ? (i + 1) % 3 : (i + 2) % 3;
            const double eps = q == 0 ? 1.0 : -1.0;
            const mat3<double> ex = ham_e(fld, V, j, s);
            const mat3<double> gp = madd(mul(ex, Ui), mul(Ui, ham_e(fld, V, j, sp)), 1.0);
            const mat3<double> gm = madd(mul(ex, W), mul(W, ham_e(fld, V, j, sm)), 1.0);
            const int mp = 1 + (2 * i + 0) * 2 + q, mm = 1 + (2 * i + 1) * 2 + q;
            for (int e = 0; e < 9; ++e) {
                blk[((int64_t)mp * 9 + e) * V + s] = {eps * inv2a * gp.m[e].re, eps * inv2a * gp.m[e].im};
                blk[((int64_t)mm * 9 + e) * V + s] = {-eps * inv2a * gm.m[e].re, -eps * inv2a * gm.m[e].im};
            }
        }
    }
    for (int e = 0; e < 9; ++e) blk[(int64_t)e * V + s] = dsum.m[e];
}

struct ham_cols {
    const cplx<double> *in[HAM_MAXCOL];
    cplx<double> *out[HAM_MAXCOL];
    cplx<double> a[HAM_MAXCOL][4];
    int n; // columns
};

// Pauli matrix entry sigma_k[r][c] of the reference's table (su2.f90:68-88):
// k = 0: [[0,1],[1,0]], k = 1: [[0,i],[-i,0]], k = 2: diag(1,-1).
__device__ __forceinline__ cplx<double> ham_sigma(int k, int r, int c)
{
    if (k == 0) return {r != c ? 1.0 : 0.0, 0.0};
    if (k == 1) return {0.0, r == c ? 0.0 : (r == 0 ? 1.0 : -1.0)};
    return {r == c ? (r == 0 ? 1.0 : -1.0) : 0.0, 0.0};
}

__device__ __forceinline__ void ham_load(const cplx<double> *__restrict__ x, int64_t V, int64_t s, cplx<double> (&v)[2][3])
{
#pragma unroll
    for (int sp = 0; sp < 2; ++sp)
#pragma unroll
        for (int c = 0; c < 3; ++c) v[sp][c] = x[((int64_t)sp * 3 + c) * V + s];
}

__device__ __forceinline__ mat3<double> ham_blk(const cplx<double> *__restrict__ blk, int64_t V, int m, int64_t s)
{
    mat3<double> r;
#pragma unroll
    for (int e = 0; e < 9; ++e) r.m[e] = blk[((int64_t)m * 9 + e) * V + s];
    return r;
}

// acc[sp] += M v[sp] (ADJ: M^+ v[sp])
template <bool ADJ>
__device__ __forceinline__ void ham_colour(const mat3<double> &M, const cplx<double> (&v)[2][3], cplx<double> (&acc)[2][3])
{
#pragma unroll
    for (int sp = 0; sp < 2; ++sp)
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (ADJ) mac_cn(acc[sp][r], M.m[3 * c + r], v[sp][c]);
                else mac(acc[sp][r], M.m[3 * r + c], v[sp][c]);
            }
}

// acc += Kron(G, sigma_k) v (ADJ: Kron(G^+, sigma_k) v; sigma is Hermitian)
template <bool ADJ>
__device__ __forceinline__ void ham_kron(const mat3<double> &G, int k, const cplx<double> (&v)[2][3], cplx<double> (&acc)[2][3])
{
    cplx<double> w[2][3];
#pragma unroll
    for (int sp = 0; sp < 2; ++sp)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            cplx<double> t = {0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 2; ++q) mac(t, ham_sigma(k, sp, q), v[q][c]);
            w[sp][c] = t;
        }
    ham_colour<ADJ>(G, w, acc);
}

// y = S x (ADJ = false) or S^+ x for the columns of `cols`.  The grid is
// site blocks x columns with the column fastest, and each XCD walks one
// contiguous range of it: the columns of a site block run on one XCD one
// after the other and share its links, field strengths and colour blocks
// (2.4 KB per site) through that XCD's L2.
template <typename R, bool ADJ>
__global__ __launch_bounds__(HAM_THREADS) void k_ham_apply(const cplx<R> *__restrict__ u, const double *__restrict__ fld,
                                                          const cplx<double> *__restrict__ blk, geom g, ham_cols cols)
{
    const unsigned lb = xcd_block(gridDim.x, 8);
    const int col = (int)(lb % (unsigned)cols.n);
    const int64_t s = (int64_t)(lb / (unsigned)cols.n) * blockDim.x + threadIdx.x;
    if (s >= g.vol) return;
    const int64_t V = g.vol;
    const cplx<double> *x = cols.in[col];
    cplx<double> a[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) a[t] = ADJ ? cplx<double>{cols.a[col][t].re, -cols.a[col][t].im} : cols.a[col][t];
    const lsite p = make_site(g, s);
    cplx<double> v0[2][3], kin[2][3], c2[2][3], c3[2][3], c4[2][3];
    ham_load(x, V, s, v0);
#pragma unroll
    for (int sp = 0; sp < 2; ++sp)
#pragma unroll
        for (int c = 0; c < 3; ++c) kin[sp][c] = c2[sp][c] = c3[sp][c] = c4[sp][c] = {0.0, 0.0};
    // c2, c3 (site-diagonal colour matrices)
    {
        mat3<double> B;
        for (int e = 0; e < 9; ++e) B.m[e] = {fld[((int64_t)27 + e) * V + s], 0.0};
        ham_colour<ADJ>(B, v0, c2);
        ham_colour<ADJ>(ham_blk(blk, V, 0, s), v0, c3);
    }
    for (int i = 0; i < 3; ++i) {
        const int64_t sp_ = site_up(g, p, i).s, sm = site_dn(g, p, i).s;
        const double w2 = 1.0 / (g.a[i + 1] * g.a[i + 1]);
        cplx<double> vp[2][3], vm[2][3], h[2][3];
        ham_load(x, V, sp_, vp);
        ham_load(x, V, sm, vm);
        // kinetic: v0 - (U_i(x) v(x+i) + U_i(x-i)^+ v(x-i)) / 2
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) h[q][c] = {0.0, 0.0};
        ham_colour<false>(ham_link(u, V, i, s), vp, h);
        ham_colour<true>(ham_link(u, V, i, sm), vm, h);
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                kin[q][c] = kin[q][c] + w2 * (v0[q][c] - 0.5 * h[q][c]);
        // spin-orbit hops
        for (int q = 0; q < 2; ++q) {
            const int k = q == 0 ? (i + 2) % 3 : (i + 1) % 3;
            const int mp = 1 + (2 * i + 0) * 2 + q, mm = 1 + (2 * i + 1) * 2 + q;
            if (!ADJ) {
                ham_kron<false>(ham_blk(blk, V, mp, s), k, vp, c4);
                ham_kron<false>(ham_blk(blk, V, mm, s), k, vm, c4);
            } else {
                ham_kron<true>(ham_blk(blk, V, mp, sm), k, vm, c4);
                ham_kron<true>(ham_blk(blk, V, mm, sp_), k, vp, c4);
            }
        }
    }
    cplx<double> *y = cols.out[col];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            cplx<double> r = v0[q][c];
            mac(r, a[0], kin[q][c]);
            mac(r, a[1], c2[q][c]);
            mac(r, a[2], c3[q][c]);
            mac(r, a[3], c4[q][c]);
            y[((int64_t)q * 3 + c) * V + s] = r;
        }
}

// ---- vector kernels (n complex doubles) -----------------------------------

__device__ __forceinline__ double ham_wsum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__device__ __forceinline__ double ham_block_sum(double v, double *sh)
{
    v = ham_wsum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += sh[k];
    return t;
}

// Per-column pointers of the Krylov vector kernels: column q of a launch
// (blockIdx.z for k_ham_dots, blockIdx.y for the others) works on basis[q],
// w[q], partial[q], small[q] with scale[q]; every column's sums are the same
// as in a launch of its own.
struct ham_vecs {
    const cplx<double> *basis[HAM_MAXCOL];
    cplx<double> *w[HAM_MAXCOL];
    double *partial[HAM_MAXCOL];
    double *small[HAM_MAXCOL];
    double scale[HAM_MAXCOL];
};

// partial[(2 i + {0: re, 1: im}) * gridDim.x + block] of <V_i, w>, i = blockIdx.y
__global__ __launch_bounds__(HAM_THREADS) void k_ham_dots(ham_vecs t, int64_t n)
{
    const int i = blockIdx.y, q = blockIdx.z;
    const cplx<double> *vi = t.basis[q] + (int64_t)i * n;
    const cplx<double> *__restrict__ w = t.w[q];
    double *__restrict__ partial = t.partial[q];
    cplx<double> acc = {0.0, 0.0};
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        mac_cn(acc, vi[e], w[e]);
    __shared__ double sh[HAM_THREADS / 64];
    const double re = ham_block_sum(acc.re, sh);
    const double im = ham_block_sum(acc.im, sh);
    if (threadIdx.x == 0) {
        partial[(int64_t)(2 * i) * gridDim.x + blockIdx.x] = re;
        partial[(int64_t)(2 * i + 1) * gridDim.x + blockIdx.x] = im;
    }
}

// partial[block] = sum |w|^2
__global__ __launch_bounds__(HAM_THREADS) void k_ham_norm2(ham_vecs t, int64_t n)
{
    const cplx<double> *__restrict__ w = t.w[blockIdx.y];
    double *__restrict__ partial = t.partial[blockIdx.y];
    double acc = 0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        acc += w[e].re * w[e].re + w[e].im * w[e].im;
    __shared__ double sh[HAM_THREADS / 64];
    const double tot = ham_block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// out[q] = sum_b partial[q * nb + b] for q = blockIdx.x (fixed order), out =
// small + offset of the column
__global__ __launch_bounds__(HAM_THREADS) void k_ham_fold(ham_vecs t, int nb, int offset)
{
    const double *__restrict__ partial = t.partial[blockIdx.y];
    double *__restrict__ out = t.small[blockIdx.y] + offset;
    const int q = blockIdx.x;
    double s = 0;
    for (int b = threadIdx.x; b < nb; b += HAM_THREADS) s += partial[(int64_t)q * nb + b];
    __shared__ double sh[HAM_THREADS / 64];
    s = ham_block_sum(s, sh);
    if (threadIdx.x == 0) out[q] = s;
}

// w += sign sum_{i < k} c_i V_i (c: complex coefficients on the device, the
// column's small)
__global__ __launch_bounds__(HAM_THREADS) void k_ham_combine(ham_vecs t, int64_t n, int k, double sign)
{
    cplx<double> *__restrict__ w = t.w[blockIdx.y];
    const cplx<double> *__restrict__ basis = t.basis[blockIdx.y];
    const double *__restrict__ c = t.small[blockIdx.y];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        cplx<double> acc = {0.0, 0.0};
        for (int i = 0; i < k; ++i) mac(acc, cplx<double>{c[2 * i], c[2 * i + 1]}, basis[(int64_t)i * n + e]);
        w[e] = {w[e].re + sign * acc.re, w[e].im + sign * acc.im};
    }
}

// w = scale w for the columns of the table
__global__ __launch_bounds__(HAM_THREADS) void k_ham_scale(ham_vecs t, int64_t n)
{
    cplx<double> *__restrict__ w = t.w[blockIdx.y];
    const double a = t.scale[blockIdx.y];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        w[e] = {a * w[e].re, a * w[e].im};
}

// y = a x + b z (real a, b)
__global__ __launch_bounds__(HAM_THREADS) void k_ham_axpby(cplx<double> *__restrict__ y, const cplx<double> *__restrict__ x,
                                                          double a, const cplx<double> *__restrict__ z, double b, int64_t n)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        y[e] = {a * x[e].re + b * z[e].re, a * x[e].im + b * z[e].im};
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct ham_workspace {
    scratch_buf fld, blk, basis, vec, red, small;
    double *host = nullptr; // pinned: 2 (m + 2) doubles, kept until the process ends
    int host_m = -1;
};
static ham_workspace g_ham;

static unsigned ham_blocks(int64_t n)
{
    return (unsigned)std::min<int64_t>(std::max<int64_t>((n + HAM_THREADS - 1) / HAM_THREADS, 1), 8192);
}

struct ham_operator {
    const void *u;
    int precision;
    geom g;
    const double *fld;
    const cplx<double> *blk;
};

static int ham_apply(const ham_operator &op, bool adj, int ncol, const cplx<double> *const *in, cplx<double> *const *out,
                     const cplx<double> (*a)[4])
{
    ham_cols cols;
    for (int c = 0; c < ncol; ++c) {
        cols.in[c] = in[c];
        cols.out[c] = out[c];
        for (int t = 0; t < 4; ++t) cols.a[c][t] = a[c][t];
    }
    cols.n = ncol;
    const dim3 grd((unsigned)((op.g.vol + HAM_THREADS - 1) / HAM_THREADS) * (unsigned)ncol);
    if (op.precision == 64) {
        if (adj)
            hipLaunchKernelGGL((k_ham_apply<double, true>), grd, dim3(HAM_THREADS), 0, stream(), (const cplx<double> *)op.u, op.fld, op.blk, op.g, cols);
        else
            hipLaunchKernelGGL((k_ham_apply<double, false>), grd, dim3(HAM_THREADS), 0, stream(), (const cplx<double> *)op.u, op.fld, op.blk, op.g, cols);
    } else {
        if (adj)
            hipLaunchKernelGGL((k_ham_apply<float, true>), grd, dim3(HAM_THREADS), 0, stream(), (const cplx<float> *)op.u, op.fld, op.blk, op.g, cols);
        else
            hipLaunchKernelGGL((k_ham_apply<float, false>), grd, dim3(HAM_THREADS), 0, stream(), (const cplx<float> *)op.u, op.fld, op.blk, op.g, cols);
    }
    NRQ_HIP(hipGetLastError());
    return NRQ_OK;
}

// GMRES(m) for S^+ x_c = b_c on nc columns (n = 6V complex each), zero
// initial guess, stopping on ||b_c - S^+ x_c|| <= tol ||b_c||.  The columns
// run in lockstep, each through exactly the operations (and summation order)
// it would take alone: one operator launch for all columns of an Arnoldi
// step and one host synchronisation per step for all their Gram-Schmidt
// coefficients.  Column c uses basis block c of the workspace.  A column
// that reaches maxit leaves NRQ_ERR_NOCONV as the result; the others finish.
static int ham_gmres(const ham_operator &op, int nc, const cplx<double> (*a)[4], cplx<double> *const *b,
                     cplx<double> *const *x, int64_t n, int m, int64_t maxit, double tol, nrq_solve_stats *st)
{
    const int64_t ldb = (int64_t)(m + 1) * n;                 // basis block per column
    const int64_t lred = 2 * (int64_t)(m + 2) * HAM_DOT_BLOCKS; // reduction partials per column
    const int lsm = 2 * (m + 2);                               // small device / pinned host slots per column
    auto basis = [&](int c) { return (cplx<double> *)g_ham.basis.ptr + c * ldb; };
    auto rvec = [&](int c) { return (cplx<double> *)g_ham.vec.ptr + (int64_t)c * n; };
    auto red = [&](int c) { return (double *)g_ham.red.ptr + c * lred; };
    auto dsm = [&](int c) { return (double *)g_ham.small.ptr + (int64_t)c * lsm; };
    auto hst = [&](int c) { return g_ham.host + (int64_t)c * lsm; };
    const unsigned nb = ham_blocks(n);
    // the Krylov-kernel table of the columns sel[0 .. ns) with vectors wv
    auto table = [&](int ns, const int *sl, cplx<double> *const *wv) {
        ham_vecs t{};
        for (int q = 0; q < ns; ++q) {
            t.basis[q] = basis(sl[q]);
            t.w[q] = wv[q];
            t.partial[q] = red(sl[q]);
            t.small[q] = dsm(sl[q]);
            t.scale[q] = 1.0;
        }
        return t;
    };
    // the small device slots of all columns to the pinned host slots
    auto smalls_to_host = [&]() {
        return hipMemcpyAsync(hst(0), dsm(0), sizeof(double) * (size_t)lsm * (size_t)nc, hipMemcpyDeviceToHost, stream());
    };
    using zc = std::complex<double>;
    struct column {
        std::vector<zc> H, cs, sn, g, y;
        double bnorm = 0, rnorm = 0;
        int64_t its = 0;
        int k = 0;           // Arnoldi steps of the current cycle
        bool active = false; // not yet converged
        bool cycling = false;
        bool failed = false;
    };
    std::vector<column> col(nc);
    const cplx<double> *ins[HAM_MAXCOL];
    cplx<double> *outs[HAM_MAXCOL];
    cplx<double> cof[HAM_MAXCOL][4];
    int sel[HAM_MAXCOL];
    // ||b||^2 of every column, one synchronisation
    for (int c = 0; c < nc; ++c) {
        sel[c] = c;
        outs[c] = b[c];
    }
    {
        const ham_vecs t = table(nc, sel, outs);
        hipLaunchKernelGGL(k_ham_norm2, dim3(HAM_DOT_BLOCKS, nc), dim3(HAM_THREADS), 0, stream(), t, n);
        hipLaunchKernelGGL(k_ham_fold, dim3(1, nc), dim3(HAM_THREADS), 0, stream(), t, HAM_DOT_BLOCKS, 0);
    }
    NRQ_HIP(smalls_to_host());
    NRQ_HIP(hipGetLastError());
    NRQ_HIP(hipStreamSynchronize(stream()));
    for (int c = 0; c < nc; ++c) {
        column &C = col[c];
        const double bn2 = hst(c)[0];
        NRQ_HIP(hipMemsetAsync(x[c], 0, sizeof(cplx<double>) * (size_t)n, stream()));
        if (bn2 == 0.0) continue; // x = 0, as KSPSolve returns for b = 0
        ++st->solves;
        C.active = true;
        C.bnorm = C.rnorm = std::sqrt(bn2);
        C.H.assign((size_t)(m + 1) * m, 0.0);
        C.cs.assign(m, 0.0);
        C.sn.assign(m, 0.0);
        C.g.assign(m + 1, 0.0);
        C.y.assign(m, 0.0);
        // r = b (x = 0)
        hipLaunchKernelGGL(k_ham_axpby, dim3(nb), dim3(HAM_THREADS), 0, stream(), rvec(c), b[c], 1.0, b[c], 0.0, n);
    }
    while (true) {
        int nact = 0;
        for (int c = 0; c < nc; ++c) nact += col[c].active;
        if (!nact) break;
        // new cycle: V_0 = r / ||r||
        for (int c = 0; c < nc; ++c) {
            column &C = col[c];
            if (!C.active) continue;
            ++st->restarts;
            hipLaunchKernelGGL(k_ham_axpby, dim3(nb), dim3(HAM_THREADS), 0, stream(), basis(c), rvec(c), 1.0 / C.rnorm,
                               rvec(c), 0.0, n);
            std::fill(C.g.begin(), C.g.end(), 0.0);
            C.g[0] = C.rnorm;
            C.k = 0;
            C.cycling = true;
        }
        for (int k = 0; k < m; ++k) {
            int ns = 0;
            for (int c = 0; c < nc; ++c) {
                column &C = col[c];
                if (C.active && C.cycling && C.its >= maxit) C.cycling = false;
                if (!(C.active && C.cycling)) continue;
                ++C.its;
                sel[ns] = c;
                ins[ns] = basis(c) + (int64_t)k * n;
                outs[ns] = basis(c) + (int64_t)(k + 1) * n;
                for (int t = 0; t < 4; ++t) cof[ns][t] = a[c][t];
                ++ns;
            }
            if (!ns) break;
            if (int e = ham_apply(op, true, ns, ins, outs, cof)) return e;
            // classical Gram-Schmidt of every column: h = V^+ w, w -= V h, ||w||^2
            ham_vecs tab = table(ns, sel, outs);
            hipLaunchKernelGGL(k_ham_dots, dim3(HAM_DOT_BLOCKS, k + 1, ns), dim3(HAM_THREADS), 0, stream(), tab, n);
            hipLaunchKernelGGL(k_ham_fold, dim3(2 * (k + 1), ns), dim3(HAM_THREADS), 0, stream(), tab, HAM_DOT_BLOCKS, 0);
            hipLaunchKernelGGL(k_ham_combine, dim3(nb, ns), dim3(HAM_THREADS), 0, stream(), tab, n, k + 1, -1.0);
            hipLaunchKernelGGL(k_ham_norm2, dim3(HAM_DOT_BLOCKS, ns), dim3(HAM_THREADS), 0, stream(), tab, n);
            hipLaunchKernelGGL(k_ham_fold, dim3(1, ns), dim3(HAM_THREADS), 0, stream(), tab, HAM_DOT_BLOCKS, 2 * (k + 1));
            NRQ_HIP(smalls_to_host());
            NRQ_HIP(hipGetLastError());
            NRQ_HIP(hipStreamSynchronize(stream()));
            int nsc = 0; // columns to normalise, in tab's first nsc slots
            for (int q = 0; q < ns; ++q) {
                const int c = sel[q];
                column &C = col[c];
                const double *host = hst(c);
                std::vector<zc> &H = C.H;
                for (int i = 0; i <= k; ++i) H[(size_t)i * m + k] = {host[2 * i], host[2 * i + 1]};
                const double hk1 = std::sqrt(host[2 * (k + 1)]);
                H[(size_t)(k + 1) * m + k] = hk1;
                if (hk1 > 0) {
                    tab.w[nsc] = outs[q];
                    tab.scale[nsc] = 1.0 / hk1;
                    ++nsc;
                }
                // Givens rotations
                for (int i = 0; i < k; ++i) {
                    const zc t = C.cs[i] * H[(size_t)i * m + k] + C.sn[i] * H[(size_t)(i + 1) * m + k];
                    H[(size_t)(i + 1) * m + k] = -std::conj(C.sn[i]) * H[(size_t)i * m + k] + C.cs[i] * H[(size_t)(i + 1) * m + k];
                    H[(size_t)i * m + k] = t;
                }
                const zc hkk = H[(size_t)k * m + k];
                const double den = std::sqrt(std::norm(hkk) + hk1 * hk1);
                if (den == 0.0) {
                    C.cs[k] = 1.0;
                    C.sn[k] = 0.0;
                } else {
                    C.cs[k] = std::abs(hkk) / den;
                    const zc ph = std::abs(hkk) > 0 ? hkk / std::abs(hkk) : zc(1.0, 0.0);
                    C.sn[k] = ph * hk1 / den;
                }
                H[(size_t)k * m + k] = C.cs[k] * hkk + C.sn[k] * hk1;
                H[(size_t)(k + 1) * m + k] = 0.0;
                C.g[k + 1] = -std::conj(C.sn[k]) * C.g[k];
                C.g[k] = C.cs[k] * C.g[k];
                C.k = k + 1;
                if (std::abs(C.g[k + 1]) <= 0.5 * tol * C.bnorm || hk1 == 0.0) C.cycling = false;
            }
            if (nsc) hipLaunchKernelGGL(k_ham_scale, dim3(nb, nsc), dim3(HAM_THREADS), 0, stream(), tab, n);
        }
        // end of the cycle: y = H^-1 g (upper triangular, k x k), x += V y, then
        // the true residual r = b - S^+ x of every column that ran
        int ns = 0;
        for (int c = 0; c < nc; ++c) {
            column &C = col[c];
            if (!C.active) continue;
            const int k = C.k;
            for (int i = k - 1; i >= 0; --i) {
                zc t = C.g[i];
                for (int j = i + 1; j < k; ++j) t -= C.H[(size_t)i * m + j] * C.y[j];
                C.y[i] = t / C.H[(size_t)i * m + i];
            }
            double *host = hst(c);
            for (int i = 0; i < k; ++i) {
                host[2 * i] = C.y[i].real();
                host[2 * i + 1] = C.y[i].imag();
            }
            NRQ_HIP(hipMemcpyAsync(dsm(c), host, sizeof(double) * 2 * k, hipMemcpyHostToDevice, stream()));
            {
                const int one[1] = {c};
                hipLaunchKernelGGL(k_ham_combine, dim3(nb, 1), dim3(HAM_THREADS), 0, stream(), table(1, one, &x[c]), n, k, 1.0);
            }
            sel[ns] = c;
            ins[ns] = x[c];
            outs[ns] = rvec(c);
            for (int t = 0; t < 4; ++t) cof[ns][t] = a[c][t];
            ++ns;
        }
        if (int e = ham_apply(op, true, ns, ins, outs, cof)) return e;
        for (int q = 0; q < ns; ++q) {
            const int c = sel[q];
            hipLaunchKernelGGL(k_ham_axpby, dim3(nb), dim3(HAM_THREADS), 0, stream(), rvec(c), b[c], 1.0, rvec(c), -1.0, n);
        }
        {
            const ham_vecs t = table(ns, sel, outs);
            hipLaunchKernelGGL(k_ham_norm2, dim3(HAM_DOT_BLOCKS, ns), dim3(HAM_THREADS), 0, stream(), t, n);
            hipLaunchKernelGGL(k_ham_fold, dim3(1, ns), dim3(HAM_THREADS), 0, stream(), t, HAM_DOT_BLOCKS, 0);
        }
        NRQ_HIP(smalls_to_host());
        NRQ_HIP(hipGetLastError());
        NRQ_HIP(hipStreamSynchronize(stream()));
        for (int q = 0; q < ns; ++q) {
            column &C = col[sel[q]];
            C.rnorm = std::sqrt(hst(sel[q])[0]);
            st->iterations += C.k;
            if (C.rnorm <= tol * C.bnorm) {
                C.active = false;
                st->max_true_relres = std::max(st->max_true_relres, C.rnorm / C.bnorm);
            } else if (C.its >= maxit) {
                C.active = false;
                C.failed = true;
                st->max_true_relres = std::max(st->max_true_relres, C.rnorm / C.bnorm);
            }
        }
    }
    for (int c = 0; c < nc; ++c)
        if (col[c].failed)
            return fail(NRQ_ERR_NOCONV, "CN (full Hamiltonian): no convergence after %lld iterations (relres %.3e)",
                        (long long)col[c].its, col[c].rnorm / col[c].bnorm);
    return NRQ_OK;
}

int ham_cn_update(nrq_heavy *const *hs, const double *mass, int n, const nrq_gauge *gf, const double c[8],
                  double stepwidth, int restart, int64_t maxit, double tol, nrq_solve_stats *stats)
{
    const geom g = lat().g;
    const int64_t V = g.vol;
    const int64_t nvec = 6 * V;
    const int m = std::max(1, restart);
    // columns solved together: as many as fit HAM_BATCH_BYTES (NRQ_HAM_BATCH_BYTES
    // in the environment overrides it) of Krylov basis, residual and right-hand
    // side (m + 3 vectors each), at most HAM_MAXCOL
    size_t budget = HAM_BATCH_BYTES;
    if (const char *e = getenv("NRQ_HAM_BATCH_BYTES")) {
        const long long v = atoll(e);
        if (v > 0) budget = (size_t)v;
    }
    const size_t per_col = sizeof(cplx<double>) * (size_t)(m + 3) * (size_t)nvec;
    const int ncb = (int)std::max<size_t>(1, std::min<size_t>(HAM_MAXCOL, budget / per_col));
    if (int r = g_ham.fld.reserve(sizeof(double) * HAM_FLD * (size_t)V)) return r;
    if (int r = g_ham.blk.reserve(sizeof(cplx<double>) * 9 * HAM_BLK * (size_t)V)) return r;
    if (int r = g_ham.basis.reserve(sizeof(cplx<double>) * (size_t)(m + 1) * (size_t)nvec * ncb)) return r;
    if (int r = g_ham.vec.reserve(sizeof(cplx<double>) * 2 * (size_t)nvec * ncb)) return r;
    if (int r = g_ham.red.reserve(sizeof(double) * 2 * (size_t)(m + 2) * HAM_DOT_BLOCKS * ncb)) return r;
    if (int r = g_ham.small.reserve(sizeof(double) * 2 * (size_t)(m + 2) * ncb)) return r;
    const int host_need = (m + 2) * ncb;
    if (g_ham.host_m < host_need) {
        if (g_ham.host) (void)hipHostFree(g_ham.host);
        g_ham.host = nullptr;
        if (hipHostMalloc((void **)&g_ham.host, sizeof(double) * 2 * (size_t)host_need, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            g_ham.host_m = -1;
            return fail(NRQ_ERR_ALLOC, "CN (full Hamiltonian): pinned buffer allocation failed");
        }
        g_ham.host_m = host_need;
    }
    ham_operator op;
    op.u = gauge_links(gf);
    op.precision = gf->precision;
    op.g = g;
    op.fld = (const double *)g_ham.fld.ptr;
    op.blk = (const cplx<double> *)g_ham.blk.ptr;
    const unsigned nbs = (unsigned)((V + HAM_THREADS - 1) / HAM_THREADS);
    if (gf->precision == 64) {
        hipLaunchKernelGGL(k_ham_fields<double>, dim3(nbs), dim3(HAM_THREADS), 0, stream(), (const cplx<double> *)op.u,
                           (const double *)gf->e, g, (double *)g_ham.fld.ptr);
        hipLaunchKernelGGL(k_ham_blocks<double>, dim3(nbs), dim3(HAM_THREADS), 0, stream(), (const cplx<double> *)op.u,
                           (const double *)g_ham.fld.ptr, g, (cplx<double> *)g_ham.blk.ptr);
    } else {
        hipLaunchKernelGGL(k_ham_fields<float>, dim3(nbs), dim3(HAM_THREADS), 0, stream(), (const cplx<float> *)op.u,
                           (const float *)gf->e, g, (double *)g_ham.fld.ptr);
        hipLaunchKernelGGL(k_ham_blocks<float>, dim3(nbs), dim3(HAM_THREADS), 0, stream(), (const cplx<float> *)op.u,
                           (const double *)g_ham.fld.ptr, g, (cplx<double> *)g_ham.blk.ptr);
    }
    NRQ_HIP(hipGetLastError());
    const double a0 = g.a[0];
    const std::complex<double> pre(0.0, -stepwidth / 2.0 * a0), I(0.0, 1.0);
    // right-hand sides after the ncb residual vectors
    auto bvec = [&](int c) { return (cplx<double> *)g_ham.vec.ptr + (int64_t)(ncb + c) * nvec; };
    nrq_solve_stats st = {0, 0, 0, 0.0};
    int status = NRQ_OK;
    for (int hk = 0; hk < n; ++hk) {
        nrq_heavy *h = hs[hk];
        // the live columns of both propagators, solved ncb at a time
        std::vector<int> f0s;
        std::vector<std::array<cplx<double>, 4>> coefs;
        for (int prop = 0; prop < 2; ++prop) {
            const double M = prop == 0 ? mass[hk] : -mass[hk];
            std::complex<double> cc[4];
            for (int t = 0; t < 4; ++t) {
                cc[t] = std::complex<double>(c[2 * t], c[2 * t + 1]);
                if (prop == 1) cc[t] = std::conj(cc[t]);
            }
            const std::complex<double> at[4] = {pre * cc[0] / M, pre * (-cc[1]) / (2.0 * M),
                                                pre * (-cc[2]) / (8.0 * M * M), pre * (-cc[3]) * I / (8.0 * M * M)};
            std::array<cplx<double>, 4> a;
            for (int t = 0; t < 4; ++t) a[t] = {at[t].real(), at[t].imag()};
            for (int j = 0; j < 6; ++j) {
                const int f0 = 12 * prop + 2 * j;
                if (!((h->live >> f0) & 3u)) continue;
                f0s.push_back(f0);
                coefs.push_back(a);
            }
        }
        for (size_t c0 = 0; c0 < f0s.size(); c0 += (size_t)ncb) {
            const int nc = (int)std::min<size_t>((size_t)ncb, f0s.size() - c0);
            const cplx<double> *xin[HAM_MAXCOL];
            cplx<double> *xcol[HAM_MAXCOL], *bcol[HAM_MAXCOL];
            cplx<double> cof[HAM_MAXCOL][4];
            for (int c = 0; c < nc; ++c) {
                xcol[c] = (cplx<double> *)h->psi + (int64_t)f0s[c0 + c] * 3 * V;
                xin[c] = xcol[c];
                bcol[c] = bvec(c);
                for (int t = 0; t < 4; ++t) cof[c][t] = coefs[c0 + c][t];
            }
            // rhs = S psi for the batch, then S^+ psi' = rhs
            if (int e = ham_apply(op, false, nc, xin, bcol, cof)) return e;
            const int r = ham_gmres(op, nc, cof, bcol, xcol, nvec, m, maxit, tol, &st);
            for (int c = 0; c < nc; ++c) h->live |= 3u << f0s[c0 + c]; // the spin-orbit term fills both spin rows
            if (r == NRQ_ERR_NOCONV) {
                status = r;
                continue;
            }
            if (r != NRQ_OK) return r;
        }
    }
    if (stats) *stats = st;
    return status;
}

} // namespace nrq
