// nrq_gauge.hip -- gauge configuration in HBM: leapfrog kernels, FP64
// observables and the gaugeconfiguration_su3 entry points of the C ABI.
//
// HBM layout (structure of arrays, one plane per matrix element so a
// wavefront's 64 consecutive sites read 64 consecutive elements):
//   links  U[dir][e][site]  complex (R re, R im), e = 3*row + col, dir 0..2
//   efield E[dir][a][site]  R,                      a = 0..7
// with site = lattice index - 1 (x fastest).  Two link buffers ping-pong so
// the fused step can read the old links of a site's neighbours while writing
// new links.
#include "nrq_internal.h"
#include "nrq_map.h"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace nrq {

template <typename R>
__device__ __forceinline__ mat3<R> load_link(const cplx<R> *__restrict__ u, int64_t vol, int dir, int64_t s)
{
    mat3<R> m;
    const cplx<R> *p = u + (int64_t)dir * 9 * vol + s;
#pragma unroll
    for (int e = 0; e < 9; ++e) m.m[e] = p[(int64_t)e * vol];
    return m;
}

template <typename R>
__device__ __forceinline__ void store_link(cplx<R> *__restrict__ u, int64_t vol, int dir, int64_t s, const mat3<R> &m)
{
    cplx<R> *p = u + (int64_t)dir * 9 * vol + s;
#pragma unroll
    for (int e = 0; e < 9; ++e) p[(int64_t)e * vol] = m.m[e];
}

// Neighbour link for the staples.  REC: only rows 1 and 2 are read and row 3
// is rebuilt as conj(row1 x row2), exact for SU(3) (unit determinant,
// unitary), which cuts the bytes of every neighbour fetch by a third; the
// stored links keep all nine entries.
template <bool REC, typename R>
__device__ __forceinline__ mat3<R> load_nb(const cplx<R> *__restrict__ u, int64_t vol, int dir, int64_t s)
{
    if (!REC) return load_link(u, vol, dir, s);
    mat3<R> m;
    const cplx<R> *p = u + (int64_t)dir * 9 * vol;
#pragma unroll
    for (int e = 0; e < 6; ++e) m.m[e] = *at32(p + (int64_t)e * vol, (uint32_t)s);
    // row3_c = conj(row1_{c+1} row2_{c+2} - row1_{c+2} row2_{c+1})
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        cplx<R> t = {R(0), R(0)};
        mac_cc(t, m.m[c1], m.m[3 + c2]);
        const cplx<R> a = {-m.m[c2].re, -m.m[c2].im};
        mac_cc(t, a, m.m[3 + c1]);
        m.m[6 + c] = t;
    }
    return m;
}

// Per-(site, dir) E and output-link access of the leapfrog kernels; NT:
// streaming (nontemporal) loads and stores, since no other thread reads them
// in the same pass.
template <bool NT, typename R> __device__ __forceinline__ R ld_e(const R *p)
{
    if (NT) return __builtin_nontemporal_load(p);
    return *p;
}
template <bool NT, typename R> __device__ __forceinline__ void st_e(R *p, R v)
{
    if (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}
template <bool NT, typename R>
__device__ __forceinline__ void store_link_out(cplx<R> *__restrict__ u, int64_t vol, int dir, int64_t s, const mat3<R> &m)
{
    cplx<R> *p = u + (int64_t)dir * 9 * vol + s;
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        if (NT) store_nt(p + (int64_t)e * vol, m.m[e]);
        else p[(int64_t)e * vol] = m.m[e];
    }
}

// Update_Links_Leapfrog at one (site, dir), gaugeconfiguration_su3.f90:
// 2133-2139: U <- exp(i dt E^a T^a) U.
template <typename R>
__device__ __forceinline__ mat3<R> evolve_link(const mat3<R> &u, const R *e, R dt)
{
    R al[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) al[a] = e[a] * dt;
    mat3<R> x = exp_i_hermitian(algebra_matrix(al));
    return mul(x, u);
}

// Projects a link onto SU(3) in FP64: normalise row 1, orthonormalise row 2
// against it, row 3 = conj(row1 x row2); the result rounded to R.
template <typename R> __device__ __forceinline__ void project_su3(mat3<R> &m)
{
    cplx<double> r0[3], r1[3];
    for (int c = 0; c < 3; ++c) {
        r0[c] = {(double)m.m[c].re, (double)m.m[c].im};
        r1[c] = {(double)m.m[3 + c].re, (double)m.m[3 + c].im};
    }
    double n0 = 0;
    for (int c = 0; c < 3; ++c) n0 += r0[c].re * r0[c].re + r0[c].im * r0[c].im;
    n0 = 1.0 / sqrt(n0);
    for (int c = 0; c < 3; ++c) r0[c] = n0 * r0[c];
    cplx<double> pr = {0, 0};
    for (int c = 0; c < 3; ++c) pr = pr + cmul(r0[c], r1[c]);
    for (int c = 0; c < 3; ++c) r1[c] = r1[c] - pr * r0[c];
    double n1 = 0;
    for (int c = 0; c < 3; ++c) n1 += r1[c].re * r1[c].re + r1[c].im * r1[c].im;
    n1 = 1.0 / sqrt(n1);
    for (int c = 0; c < 3; ++c) r1[c] = n1 * r1[c];
    cplx<double> r2[3];
    r2[0] = conjg(r0[1] * r1[2] - r0[2] * r1[1]);
    r2[1] = conjg(r0[2] * r1[0] - r0[0] * r1[2]);
    r2[2] = conjg(r0[0] * r1[1] - r0[1] * r1[0]);
    for (int c = 0; c < 3; ++c) {
        m.m[c] = {(R)r0[c].re, (R)r0[c].im};
        m.m[3 + c] = {(R)r1[c].re, (R)r1[c].im};
        m.m[6 + c] = {(R)r2[c].re, (R)r2[c].im};
    }
}

// Update_Efield_Leapfrog_atSite_Direction, gaugeconfiguration_su3.f90:
// 2164-2198: E^a_i(x) -= 2 Im Tr[T^a U_i(x) sum_{k!=i} dt (a0/a_k)^2 (S^U+S^D)]
// with S^U = U_k(x+i) U_i(x+k)^+ U_k(x)^+ and
//      S^D = U_k(x+i-k)^+ U_i(x-k)^+ U_k(x-k).
// staple_sum returns sum_{k!=i} dt (a0/a_k)^2 (S^U+S^D); apply_force does the
// rest.
template <bool REC, typename R>
__device__ __forceinline__ mat3<R> staple_sum(const cplx<R> *__restrict__ u, const geom &g, int i, const lsite &p, R dt)
{
    mat3<R> sum;
    set_zero(sum);
    // k runs in increasing order, as in the reference's staple sum
    for (int k = 0; k < 3; ++k) {
        if (k == i) continue;
        const R coef = dt * R((g.a[0] / g.a[k + 1]) * (g.a[0] / g.a[k + 1]));
        const lsite pmk = site_dn(g, p, k);
        const int64_t xi = site_up(g, p, i).s;
        const int64_t xk = site_up(g, p, k).s;
        const int64_t xmk = pmk.s;
        const int64_t ximk = site_up(g, pmk, i).s;
        const mat3<R> ss = staple_pair(load_nb<REC>(u, g.vol, k, xi), load_nb<REC>(u, g.vol, i, xk),
                                       load_nb<REC>(u, g.vol, k, p.s), load_nb<REC>(u, g.vol, k, ximk),
                                       load_nb<REC>(u, g.vol, i, xmk), load_nb<REC>(u, g.vol, k, xmk));
#pragma unroll
        for (int m = 0; m < 9; ++m) sum.m[m] = sum.m[m] + coef * ss.m[m];
    }
    return sum;
}

template <typename R> __device__ __forceinline__ void apply_force(const mat3<R> &ui, const mat3<R> &sum, R *e)
{
    mat3<R> w = mul(ui, sum);
    R t[8];
    two_im_trace_gen(w, t);
#pragma unroll
    for (int a = 0; a < 8; ++a) e[a] -= t[a];
}

// E_i(x) and U_i(x) of the E update, read before the staples or, with LATE,
// after them (with 32-bit offsets), so that they hold no registers while the
// neighbour links stream in.  Both orders give the same arithmetic; which one
// schedules better depends on the neighbour format, so the two-row variants
// use LATE.
template <bool REC, bool NT, typename R>
__device__ __forceinline__ void efield_update(const cplx<R> *__restrict__ u, R *__restrict__ ef, const geom &g, int i,
                                              const lsite &p, R dt, R *e, mat3<R> &ui)
{
    const int64_t s = p.s;
    if (REC) {
        const mat3<R> sum = staple_sum<REC>(u, g, i, p, dt);
#pragma unroll
        for (int a = 0; a < 8; ++a) e[a] = ld_e<NT>(at32(ef + ((int64_t)i * 8 + a) * g.vol, (uint32_t)s));
        const cplx<R> *b = u + (int64_t)i * 9 * g.vol;
#pragma unroll
        for (int k = 0; k < 9; ++k) ui.m[k] = *at32(b + (int64_t)k * g.vol, (uint32_t)s);
        apply_force(ui, sum, e);
    } else {
#pragma unroll
        for (int a = 0; a < 8; ++a) e[a] = ld_e<NT>(ef + ((int64_t)i * 8 + a) * g.vol + s);
        ui = load_link(u, g.vol, i, s);
        apply_force(ui, staple_sum<REC>(u, g, i, p, dt), e);
    }
}

// Buffers of the configurations of one launch (blockIdx.y selects one): all
// share precision and lattice; passed to the kernels by value.
constexpr int LF_MAXB = 32;
template <typename R> struct lf_batch {
    const cplx<R> *uin[LF_MAXB];
    cplx<R> *uout[LF_MAXB];
    R *ef[LF_MAXB];
};

template <typename R, int BS, bool NT>
__global__ __launch_bounds__(BS) void k_links(lf_batch<R> bt, geom g, R dt, lf_map m)
{
    const cplx<R> *__restrict__ uin = bt.uin[blockIdx.y];
    cplx<R> *__restrict__ uout = bt.uout[blockIdx.y];
    const R *__restrict__ ef = bt.ef[blockIdx.y];
    lsite p;
    int i;
    if (!lf_site(m, g, p, i)) return;
    const int64_t s = p.s;
    const int64_t vol = g.vol;
    R e[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) e[a] = ld_e<NT>(ef + ((int64_t)i * 8 + a) * vol + s);
    mat3<R> u = load_link(uin, vol, i, s);
    store_link_out<NT>(uout, vol, i, s, evolve_link(u, e, dt));
}

template <typename R, int BS, bool NT, bool REC>
__global__ __launch_bounds__(BS) void k_efield(lf_batch<R> bt, geom g, R dt, lf_map m)
{
    const cplx<R> *__restrict__ u = bt.uin[blockIdx.y];
    R *__restrict__ ef = bt.ef[blockIdx.y];
    lsite p;
    int i;
    if (!lf_site(m, g, p, i)) return;
    const int64_t s = p.s;
    R e[8];
    mat3<R> ui;
    efield_update<REC, NT>(u, ef, g, i, p, dt, e, ui);
#pragma unroll
    for (int a = 0; a < 8; ++a) st_e<NT>(ef + ((int64_t)i * 8 + a) * g.vol + s, e[a]);
}

// One fused pass: the E update of a leapfrog step followed by the link update
// of the next step at the same (site, dir).  Both halves are exactly the
// per-(site, dir) operations of the reference; fusing them only removes one
// HBM round trip of E and U.  Neighbour links are read from uin, new links
// go to uout (ping-pong), E is updated in place (each thread owns its E).
// With reu the link is projected onto SU(3) (project_su3) after the E update
// has used it and before the link update: the same result as a separate
// reunitarisation pass between the two halves.
template <typename R, int BS, bool NT, bool REC>
__global__ __launch_bounds__(BS) void k_fused(lf_batch<R> bt, geom g, R dt, lf_map m, int reu)
{
    const cplx<R> *__restrict__ uin = bt.uin[blockIdx.y];
    cplx<R> *__restrict__ uout = bt.uout[blockIdx.y];
    R *__restrict__ ef = bt.ef[blockIdx.y];
    lsite p;
    int i;
    if (!lf_site(m, g, p, i)) return;
    const int64_t s = p.s;
    R e[8];
    mat3<R> ui;
    efield_update<REC, NT>(uin, ef, g, i, p, dt, e, ui);
#pragma unroll
    for (int a = 0; a < 8; ++a) st_e<NT>(ef + ((int64_t)i * 8 + a) * g.vol + s, e[a]);
    if (reu) project_su3(ui);
    store_link_out<NT>(uout, g.vol, i, s, evolve_link(ui, e, dt));
}

// The fused pass with the neighbour links staged in LDS.  A block owns a
// TX x TY tile of (x, y) and ZC consecutive z planes of one configuration
// (blockIdx.y); thread (site of the tile, direction i) works on (x, y, z, i)
// for z in its run.  LDS holds a ring of three z planes of links, each the
// tile with a one-site halo in x and y, all three directions and nine
// elements ([slot][dir * 9 + e][halo site]); planes z - 1, z, z + 1 serve
// every staple of plane z (the 13 links of (x, i) lie in them).  The next
// plane streams into registers while plane z is computed and replaces plane
// z - 1 afterwards, so each link is read from HBM once per run of ZC planes
// (plus the x/y halo, from L2, and two planes per run).  Per (site, dir) the arithmetic is that
// of k_fused with full neighbour links (reu as there).  The lattice extents
// must be multiples of TX, TY and ZC.
template <typename R, int TX, int TY, int ZC, bool NT>
__global__ __launch_bounds__(TX *TY * 3) void k_fused_tile(lf_batch<R> bt, geom g, R dt, int xcd, int reu)
{
    constexpr int HX = TX + 2, HY = TY + 2, HS = HX * HY;
    constexpr int NTH = TX * TY * 3;
    constexpr int PL = 27 * HS;                 // complex numbers per plane
    constexpr int PER = (PL + NTH - 1) / NTH;   // per thread and plane
    __shared__ cplx<R> lds[3 * PL];
    const cplx<R> *__restrict__ uin = bt.uin[blockIdx.y];
    cplx<R> *__restrict__ uout = bt.uout[blockIdx.y];
    R *__restrict__ ef = bt.ef[blockIdx.y];
    const int ntx = g.n0 / TX, nty = g.n1 / TY;
    const unsigned total = (unsigned)ntx * (unsigned)nty * (unsigned)(g.n2 / ZC);
    const unsigned lb = xcd_block(total, xcd);
    const int bx = (int)(lb % (unsigned)ntx), by = (int)((lb / (unsigned)ntx) % (unsigned)nty);
    const int bz = (int)(lb / ((unsigned)ntx * (unsigned)nty));
    const int x0 = bx * TX, y0 = by * TY, z0 = bz * ZC;
    const int t = threadIdx.x;
    const int64_t vol = g.vol;
    const int64_t plane = (int64_t)g.n0 * g.n1;
    auto wrapz = [&](int z) { return z < 0 ? z + g.n2 : (z >= g.n2 ? z - g.n2 : z); };

    cplx<R> pre[PER];
    auto gload = [&](int z) {
        const int zz = wrapz(z);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int q = t + k * NTH;
            if (q < PL) {
                const int hs = q % HS, de = q / HS;
                const int hx = hs % HX, hy = hs / HX;
                int gx = x0 - 1 + hx, gy = y0 - 1 + hy;
                gx = gx < 0 ? gx + g.n0 : (gx >= g.n0 ? gx - g.n0 : gx);
                gy = gy < 0 ? gy + g.n1 : (gy >= g.n1 ? gy - g.n1 : gy);
                pre[k] = uin[(int64_t)de * vol + gx + (int64_t)g.n0 * gy + plane * zz];
            }
        }
    };
    auto sstore = [&](int slot) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int q = t + k * NTH;
            if (q < PL) lds[slot * PL + q] = pre[k];
        }
    };
    for (int p = 0; p < 3; ++p) {
        gload(z0 - 1 + p);
        sstore(p);
    }
    __syncthreads();

    // TX * TY is a multiple of 64, so the direction is the same for a whole
    // wavefront: made scalar, every offset derived from it is scalar too
    static_assert((TX * TY) % 64 == 0, "a wavefront must hold one direction");
    const int i = __builtin_amdgcn_readfirstlane(t / (TX * TY));
    const int r = t - i * (TX * TY);
    const int tx = r % TX, ty = r / TX;
    // link of direction d at offset (dx, dy) in plane z + pz (pz = -1, 0, 1)
    int slot_of[3];
    auto lnk = [&](int pz, int d, int dx, int dy) {
        mat3<R> m;
        const cplx<R> *b = lds + slot_of[pz + 1] * PL + d * 9 * HS + (tx + 1 + dx) + HX * (ty + 1 + dy);
#pragma unroll
        for (int e = 0; e < 9; ++e) m.m[e] = b[e * HS];
        return m;
    };
    for (int kz = 0; kz < ZC; ++kz) {
        const int z = z0 + kz;
        const bool more = kz + 1 < ZC;
        if (more) gload(z + 2);
        slot_of[0] = kz % 3;
        slot_of[1] = (kz + 1) % 3;
        slot_of[2] = (kz + 2) % 3;
        // staples, k in increasing order as in the reference's sum
        mat3<R> sum;
        set_zero(sum);
        for (int k = 0; k < 3; ++k) {
            if (k == i) continue;
            const R coef = dt * R((g.a[0] / g.a[k + 1]) * (g.a[0] / g.a[k + 1]));
            const int ix = i == 0, iy = i == 1, kx = k == 0, ky = k == 1;
            const int pi = i == 2 ? 1 : 0;                   // plane of x + i
            const int pk = k == 2 ? 1 : 0;                   // plane of x + k
            const int pmk = k == 2 ? -1 : 0;                 // plane of x - k
            const int pik = (i == 2 ? 1 : 0) + (k == 2 ? -1 : 0); // plane of x + i - k
            const mat3<R> ss = staple_pair(lnk(pi, k, ix, iy), lnk(pk, i, kx, ky), lnk(0, k, 0, 0),
                                           lnk(pik, k, ix - kx, iy - ky), lnk(pmk, i, -kx, -ky), lnk(pmk, k, -kx, -ky));
#pragma unroll
            for (int m = 0; m < 9; ++m) sum.m[m] = sum.m[m] + coef * ss.m[m];
        }
        const int64_t s = (x0 + tx) + (int64_t)g.n0 * (y0 + ty) + plane * z;
        R e[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) e[a] = ld_e<NT>(ef + ((int64_t)i * 8 + a) * vol + s);
        mat3<R> ui = lnk(0, i, 0, 0);
        apply_force(ui, sum, e);
#pragma unroll
        for (int a = 0; a < 8; ++a) st_e<NT>(ef + ((int64_t)i * 8 + a) * vol + s, e[a]);
        if (reu) project_su3(ui);
        store_link_out<NT>(uout, vol, i, s, evolve_link(ui, e, dt));
        __syncthreads(); // the oldest plane is no longer read
        if (more) sstore(kz % 3);
        __syncthreads();
    }
}

template <typename R> __global__ void k_cold(cplx<R> *u, R *ef, int64_t vol)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= vol) return;
    for (int i = 0; i < 3; ++i) {
        for (int e = 0; e < 9; ++e) u[((int64_t)i * 9 + e) * vol + s] = {R((e % 4) == 0 ? 1 : 0), R(0)};
        for (int a = 0; a < 8; ++a) ef[((int64_t)i * 8 + a) * vol + s] = R(0);
    }
}

// Block reduction over a 64-wide wavefront then across the block's waves.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    return v;
}

constexpr int NOBS = 5; // E-field energy, magnetic energy, sum Re Tr P/3, gauss_ref sum, gauss max

// Per-site observables in FP64 (gaugeconfiguration_su3.f90:1480-1531,
// 724-813, 1769-1817), partials[q][block].  A block is OBS_SITES consecutive
// sites times the three directions: thread (lane, i) takes direction i of
// site base + lane, i.e. its electric energy and reference Gauss term, its
// part of G(x) and one plaquette plane ((0,1), (1,2), (0,2) for i = 0, 1, 2);
// the three parts of G(x) meet in LDS and are added in direction order
// before the maximum is taken.  One thread per site would hold three
// directions' links at once (over 256 registers, one wave per SIMD).
constexpr int OBS_SITES = 64;
constexpr int OBS_THREADS = 3 * OBS_SITES;
template <typename R>
__global__ __launch_bounds__(OBS_THREADS) void k_obs(const cplx<R> *__restrict__ u, const R *__restrict__ ef, geom g,
                                                     double *__restrict__ partials)
{
    double acc[NOBS] = {0, 0, 0, 0, 0};
    const double a0 = g.a[0];
    const int lane = threadIdx.x & 63, i = threadIdx.x >> 6;
    __shared__ cplx<double> gpart[3][9][OBS_SITES];
    const int pi = i == 2 ? 0 : i, pj = i == 0 ? 1 : 2; // plaquette plane of this thread
    for (int64_t base = (int64_t)blockIdx.x * OBS_SITES; base < g.vol; base += (int64_t)gridDim.x * OBS_SITES) {
        const int64_t s = base + lane;
        const bool in = s < g.vol;
        if (in) {
            const lsite p = make_site(g, s);
            const double ai = g.a[i + 1];
            double ed[8], en[8];
            const int64_t sm = site_dn(g, p, i).s;
            double esq = 0;
            for (int a = 0; a < 8; ++a) {
                ed[a] = (double)ef[((int64_t)i * 8 + a) * g.vol + s];
                en[a] = (double)ef[((int64_t)i * 8 + a) * g.vol + sm];
                const double v = ed[a] / ai / a0;
                esq += v * v;
            }
            acc[0] += esq / 2.0;
            mat3<double> um;
            {
                mat3<R> t = load_link(u, g.vol, i, sm);
                for (int m = 0; m < 9; ++m) um.m[m] = {(double)t.m[m].re, (double)t.m[m].im};
            }
            // M = E_i(x) - U_i(x-i)^+ E_i(x-i) U_i(x-i)   (algebra matrices)
            mat3<double> emat = algebra_matrix(ed);
            mat3<double> t = mul(mul_dn(um, algebra_matrix(en)), um);
            mat3<double> dm;
            for (int m = 0; m < 9; ++m) dm.m[m] = emat.m[m] - t.m[m];
            // reference Gauss deviation: D = M / a_i^2 / a0, sum_a 2|Im Tr(T^a D)|
            {
                mat3<double> d;
                const double sc = 1.0 / (ai * ai) / a0;
                for (int m = 0; m < 9; ++m) d.m[m] = sc * dm.m[m];
                double tt[8];
                two_im_trace_gen(d, tt);
                for (int a = 0; a < 8; ++a) acc[3] += fabs(tt[a]);
            }
            for (int m = 0; m < 9; ++m) gpart[i][m][lane] = (1.0 / ai) * dm.m[m];
            // magnetic part: plaquette P_{pi pj}
            const int64_t xi = site_up(g, p, pi).s, xj = site_up(g, p, pj).s;
            mat3<double> l1, l2, l3, l4;
            {
                mat3<R> a1 = load_link(u, g.vol, pi, s), a2 = load_link(u, g.vol, pj, xi);
                mat3<R> a3 = load_link(u, g.vol, pi, xj), a4 = load_link(u, g.vol, pj, s);
                for (int m = 0; m < 9; ++m) {
                    l1.m[m] = {(double)a1.m[m].re, (double)a1.m[m].im};
                    l2.m[m] = {(double)a2.m[m].re, (double)a2.m[m].im};
                    l3.m[m] = {(double)a3.m[m].re, (double)a3.m[m].im};
                    l4.m[m] = {(double)a4.m[m].re, (double)a4.m[m].im};
                }
            }
            mat3<double> pl = mul_nd(mul_nd(mul(l1, l2), l3), l4);
            const double tr = trace(pl).re;
            acc[1] += 2.0 * 3.0 * (1.0 - tr / 3.0) / g.a[pi + 1] / g.a[pj + 1];
            acc[2] += tr / 3.0;
        }
        __syncthreads();
        if (in && i == 0) {
            mat3<double> gsum;
            for (int m = 0; m < 9; ++m) {
                const cplx<double> v = gpart[0][m][lane] + gpart[1][m][lane];
                gsum.m[m] = (1.0 / a0) * (v + gpart[2][m][lane]);
            }
            double tt[8];
            two_re_trace_gen(gsum, tt);
            for (int a = 0; a < 8; ++a) acc[4] = fmax(acc[4], fabs(tt[a]));
        }
        __syncthreads();
    }
    __shared__ double red[NOBS][OBS_THREADS / 64];
    const int wv = threadIdx.x >> 6;
    for (int q = 0; q < NOBS; ++q) {
        double v = q == 4 ? wave_max(acc[q]) : wave_sum(acc[q]);
        if (lane == 0) red[q][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < NOBS) {
        const int q = threadIdx.x;
        double v = red[q][0];
        for (int w = 1; w < OBS_THREADS / 64; ++w) v = q == 4 ? fmax(v, red[q][w]) : v + red[q][w];
        partials[(int64_t)q * gridDim.x + blockIdx.x] = v;
    }
}

// Second stage of the deterministic reductions: block q (REDUCE_THREADS
// threads) folds the nblocks partials of quantity q (a maximum for
// q == max_q, else a sum) in a fixed order: strided per-thread sums, then the
// waves, then the waves' results in order.
__global__ __launch_bounds__(REDUCE_THREADS) void k_fold(const double *__restrict__ partials, int nblocks, int nq,
                                                         int max_q, double *__restrict__ out)
{
    const int q = blockIdx.x;
    if (q >= nq) return;
    const bool mx = q == max_q;
    double v = mx ? -INFINITY : 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
        const double w = partials[(int64_t)q * nblocks + b];
        v = mx ? fmax(v, w) : v + w;
    }
    v = mx ? wave_max(v) : wave_sum(v);
    __shared__ double red[REDUCE_THREADS / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) t = mx ? fmax(t, red[w]) : t + red[w];
        out[q] = t;
    }
}

// Projects U onto SU(3) in FP64: normalise row 1, orthonormalise row 2
// against it, row 3 = conj(row1 x row2).
template <typename R> __global__ void k_reunit(cplx<R> *u, int64_t vol)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (s >= vol) return;
    mat3<R> m = load_link(u, vol, i, s);
    project_su3(m);
    store_link(u, vol, i, s, m);
}

// Gather one link (possibly daggered / identity) into a 9-complex double slot.
template <typename R> __global__ void k_get_link(const cplx<R> *u, int64_t vol, int dir, int64_t s, double *out)
{
    const int e = threadIdx.x;
    if (e >= 9) return;
    if (dir == 0) {
        out[2 * e] = (e % 4) == 0 ? 1.0 : 0.0;
        out[2 * e + 1] = 0.0;
        return;
    }
    const int r = e / 3, c = e % 3;
    const int d = (dir > 0 ? dir : -dir) - 1;
    cplx<R> v = dir > 0 ? u[((int64_t)d * 9 + 3 * r + c) * vol + s] : conjg(u[((int64_t)d * 9 + 3 * c + r) * vol + s]);
    // column-major output (Fortran layout)
    out[2 * (r + 3 * c)] = (double)v.re;
    out[2 * (r + 3 * c) + 1] = (double)v.im;
}

template <typename R> __global__ void k_plaq(const cplx<R> *u, geom g, int i, int j, int64_t s, double *out)
{
    if (threadIdx.x != 0) return;
    const lsite ps = make_site(g, s);
    const int64_t xi = site_up(g, ps, i).s, xj = site_up(g, ps, j).s;
    mat3<double> l[4];
    mat3<R> a[4] = {load_link(u, g.vol, i, s), load_link(u, g.vol, j, xi), load_link(u, g.vol, i, xj), load_link(u, g.vol, j, s)};
    for (int k = 0; k < 4; ++k)
        for (int m = 0; m < 9; ++m) l[k].m[m] = {(double)a[k].m[m].re, (double)a[k].m[m].im};
    mat3<double> p = mul_nd(mul_nd(mul(l[0], l[1]), l[2]), l[3]);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            out[2 * (r + 3 * c)] = p.m[3 * r + c].re;
            out[2 * (r + 3 * c) + 1] = p.m[3 * r + c].im;
        }
}

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------
static size_t link_bytes(const nrq_gauge *h)
{
    return (size_t)3 * 9 * (size_t)h->vol * (h->precision == 64 ? 16 : 8);
}
static size_t efield_bytes(const nrq_gauge *h)
{
    return (size_t)3 * 8 * (size_t)h->vol * (h->precision == 64 ? 8 : 4);
}

static void release(nrq_gauge *h)
{
    for (int k = 0; k < 2; ++k)
        if (h->u[k]) (void)hipFree(h->u[k]);
    if (h->e) (void)hipFree(h->e);
    if (h->partials) (void)hipFree(h->partials);
    if (h->result) (void)hipHostFree(h->result);
    if (h->dlog) (void)hipFree(h->dlog);
    h->u[0] = h->u[1] = h->e = h->partials = nullptr;
    h->result = nullptr;
    h->dlog = nullptr;
    h->log_cap = h->log_n = 0;
}

static int allocate(nrq_gauge *h, int precision)
{
    release(h);
    h->precision = precision;
    h->vol = lat().g.vol;
    h->cur = 0;
    h->steps_since_reunit = 0;
    hipError_t e1 = hipMalloc(&h->u[0], link_bytes(h));
    hipError_t e2 = hipMalloc(&h->u[1], link_bytes(h));
    hipError_t e3 = hipMalloc(&h->e, efield_bytes(h));
    hipError_t e4 = hipMalloc(&h->partials, sizeof(double) * NOBS * REDUCE_BLOCKS + 64 * sizeof(double));
    hipError_t e5 = hipHostMalloc((void **)&h->result, 64 * sizeof(double), hipHostMallocDefault);
    if (e1 || e2 || e3 || e4 || e5) {
        release(h);
        return fail(NRQ_ERR_ALLOC, "gauge: device allocation of %zu bytes failed", 2 * link_bytes(h) + efield_bytes(h));
    }
    return NRQ_OK;
}

// The observable sums of the current configuration into dres (device, NOBS
// doubles), enqueued on the stream.
static int gauge_obs_enqueue(const nrq_gauge *h, double *dres)
{
    double *part = (double *)h->partials;
    const geom g = lat().g;
    if (h->precision == 64)
        hipLaunchKernelGGL(k_obs<double>, dim3(REDUCE_BLOCKS), dim3(OBS_THREADS), 0, stream(),
                           (const cplx<double> *)h->u[h->cur], (const double *)h->e, g, part);
    else
        hipLaunchKernelGGL(k_obs<float>, dim3(REDUCE_BLOCKS), dim3(OBS_THREADS), 0, stream(),
                           (const cplx<float> *)h->u[h->cur], (const float *)h->e, g, part);
    NRQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_fold, dim3(NOBS), dim3(REDUCE_THREADS), 0, stream(), part, REDUCE_BLOCKS, NOBS, 4, dres);
    NRQ_HIP(hipGetLastError());
    return NRQ_OK;
}

// energy, gauss_ref, gauss_max, plaquette from the raw sums
static void obs_row(const double *o, int64_t vol, double out[4])
{
    out[0] = o[0] + o[1];
    out[1] = o[3];
    out[2] = o[4];
    out[3] = o[2] / (3.0 * (double)vol);
}

int gauge_obs(const nrq_gauge *h, double out[NOBS])
{
    double *dres = (double *)h->partials + NOBS * REDUCE_BLOCKS;
    if (int r = gauge_obs_enqueue(h, dres)) return r;
    NRQ_HIP(hipMemcpyAsync(h->result, dres, NOBS * sizeof(double), hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    for (int q = 0; q < NOBS; ++q) out[q] = h->result[q];
    return NRQ_OK;
}

// Launch shapes the autotuner chooses from (block size, XCD remap, tiling,
// streaming, two-row neighbour links).  All variants compute the same
// update per (site, dir); the two-row form rebuilds the third row of each
// neighbour link, so results agree up to rounding.
struct lf_variant {
    int bs;
    int xcd;
    int mode;
    int nt;  // streaming loads/stores of E and of the new links
    int rec; // neighbour links read as two rows (third rebuilt)
};
// Modes 3 and 4 are the LDS-tiled fused pass (k_fused_tile, FP32, 16 x 8
// tiles, three planes in LDS, runs of 16 or 8 z planes); bs is the tile's
// thread count.  Their first and last half steps use the consecutive-site
// map.
static const lf_variant LF_VARIANTS[] = {
    {256, 8, 1, 1, 1}, {128, 8, 1, 1, 1}, {64, 8, 1, 1, 1}, {256, 8, 0, 1, 1}, {128, 8, 0, 1, 1}, {64, 8, 0, 1, 1},
    {256, 8, 2, 1, 1}, {256, 8, 2, 0, 1}, {256, 8, 0, 0, 1}, {64, 8, 0, 0, 1},  {256, 8, 1, 1, 0}, {256, 8, 0, 0, 0},
    {256, 8, 2, 0, 0}, {64, 8, 0, 0, 0},  {256, 1, 0, 0, 1}, {384, 8, 3, 0, 0}, {384, 8, 3, 1, 0}, {384, 8, 4, 0, 0},
    {384, 8, 4, 1, 0}};
constexpr int LF_NVARIANTS = sizeof(LF_VARIANTS) / sizeof(LF_VARIANTS[0]);
constexpr int LF_DEFAULT = 11; // {256, 8, 0, 0, 0}: used without tuning, valid on every lattice

struct lf_choice {
    int precision;
    int n[3]; // lattice extents the choice was tuned on
    int variant;
};
static std::vector<lf_choice> g_lf_choice;
static int g_autotune = -1; // -1: unset (env NRQ_AUTOTUNE, default on)

// Whether shape `variant` applies to lattice g at storage precision prec
// (64 or 32).  The LDS tiles are FP32 only: a 16 x 4 FP64 tile fits LDS, but
// at one wavefront per SIMD its prefetch registers do not fit.
static bool variant_ok(int variant, const geom &g, int prec)
{
    const lf_variant v = LF_VARIANTS[variant];
    if (v.mode == 1) return g.n0 <= v.bs;
    if (v.mode == 2) return g.n0 >= 16;
    if (v.mode >= 3)
        return prec == 32 && g.n0 % 16 == 0 && g.n1 % 8 == 0 && g.n2 % (v.mode == 3 ? 16 : 8) == 0;
    return true;
}

// Runs `launch` with NTV, RECV = the streaming and two-row flags of `variant`.
#define LF_NT(launch)                                                      \
    do {                                                                   \
        const lf_variant lv_ = LF_VARIANTS[variant];                       \
        if (lv_.nt && lv_.rec) {                                           \
            constexpr bool NTV = true, RECV = true; (void)RECV;            \
            launch;                                                        \
        } else if (lv_.nt) {                                               \
            constexpr bool NTV = true, RECV = false; (void)RECV;           \
            launch;                                                        \
        } else if (lv_.rec) {                                              \
            constexpr bool NTV = false, RECV = true; (void)RECV;           \
            launch;                                                        \
        } else {                                                           \
            constexpr bool NTV = false, RECV = false; (void)RECV;          \
            launch;                                                        \
        }                                                                  \
    } while (0)

static lf_map make_map(int variant, const geom &g)
{
    const lf_variant v = LF_VARIANTS[variant];
    if (v.mode >= 3) return make_lf_map(g, 256, v.xcd, 0, 3);
    return make_lf_map(g, v.bs, v.xcd, v.mode, 3);
}

template <typename R, int ZC, bool NTV>
static void launch_tile(const lf_batch<R> &bt, int nb, const geom &g, R dt, int xcd, int reu)
{
    constexpr int TX = 16, TY = 8;
    const unsigned total = (unsigned)(g.n0 / TX) * (unsigned)(g.n1 / TY) * (unsigned)(g.n2 / ZC);
    hipLaunchKernelGGL((k_fused_tile<R, TX, TY, ZC, NTV>), dim3(total, (unsigned)nb), dim3(TX * TY * 3), 0, stream(), bt,
                       g, dt, xcd, reu);
}

template <typename R> static void launch_links(int variant, const lf_batch<R> &bt, int nb, const geom &g, R dt)
{
    const lf_map m = make_map(variant, g);
    const dim3 grd(lf_blocks(m), (unsigned)nb);
    switch (m.bs) {
    case 64: LF_NT(hipLaunchKernelGGL((k_links<R, 64, NTV>), grd, dim3(64), 0, stream(), bt, g, dt, m)); break;
    case 128: LF_NT(hipLaunchKernelGGL((k_links<R, 128, NTV>), grd, dim3(128), 0, stream(), bt, g, dt, m)); break;
    default: LF_NT(hipLaunchKernelGGL((k_links<R, 256, NTV>), grd, dim3(256), 0, stream(), bt, g, dt, m)); break;
    }
}

template <typename R> static void launch_efield(int variant, const lf_batch<R> &bt, int nb, const geom &g, R dt)
{
    const lf_map m = make_map(variant, g);
    const dim3 grd(lf_blocks(m), (unsigned)nb);
    switch (m.bs) {
    case 64: LF_NT(hipLaunchKernelGGL((k_efield<R, 64, NTV, RECV>), grd, dim3(64), 0, stream(), bt, g, dt, m)); break;
    case 128: LF_NT(hipLaunchKernelGGL((k_efield<R, 128, NTV, RECV>), grd, dim3(128), 0, stream(), bt, g, dt, m)); break;
    default: LF_NT(hipLaunchKernelGGL((k_efield<R, 256, NTV, RECV>), grd, dim3(256), 0, stream(), bt, g, dt, m)); break;
    }
}

// reu: project each link onto SU(3) between the E and the link update.
template <typename R>
static void launch_fused(int variant, const lf_batch<R> &bt, int nb, const geom &g, R dt, int reu = 0)
{
    const lf_variant tv = LF_VARIANTS[variant];
    if constexpr (sizeof(R) == 4) if (tv.mode >= 3) {
        if (tv.mode == 3)
            tv.nt ? launch_tile<R, 16, true>(bt, nb, g, dt, tv.xcd, reu) : launch_tile<R, 16, false>(bt, nb, g, dt, tv.xcd, reu);
        else
            tv.nt ? launch_tile<R, 8, true>(bt, nb, g, dt, tv.xcd, reu) : launch_tile<R, 8, false>(bt, nb, g, dt, tv.xcd, reu);
        return;
    }
    const lf_map m = make_map(variant, g);
    const dim3 grd(lf_blocks(m), (unsigned)nb);
    switch (m.bs) {
    case 64: LF_NT(hipLaunchKernelGGL((k_fused<R, 64, NTV, RECV>), grd, dim3(64), 0, stream(), bt, g, dt, m, reu)); break;
    case 128: LF_NT(hipLaunchKernelGGL((k_fused<R, 128, NTV, RECV>), grd, dim3(128), 0, stream(), bt, g, dt, m, reu)); break;
    default: LF_NT(hipLaunchKernelGGL((k_fused<R, 256, NTV, RECV>), grd, dim3(256), 0, stream(), bt, g, dt, m, reu)); break;
    }
}

// one-configuration table
template <typename R> static lf_batch<R> one(const cplx<R> *uin, cplx<R> *uout, R *ef)
{
    lf_batch<R> b = {};
    b.uin[0] = uin;
    b.uout[0] = uout;
    b.ef[0] = ef;
    return b;
}

// Largest |a[k] - b[k]| over n reals, per block into part[blockIdx.x].
template <typename R>
__global__ __launch_bounds__(REDUCE_THREADS) void k_maxdiff(const R *__restrict__ a, const R *__restrict__ b, int64_t n,
                                                            double *__restrict__ part)
{
    double m = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs((double)a[k] - (double)b[k]));
    m = wave_max(m);
    __shared__ double red[REDUCE_THREADS / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
        for (int w = 1; w < REDUCE_THREADS / 64; ++w) t = fmax(t, red[w]);
        part[blockIdx.x] = t;
    }
}

// Agreement a shape must show with LF_DEFAULT after one fused pass from the
// same state: links (entries of modulus <= 1) to `tol`, E to tol * max(1, |E|).
// Two-row shapes rebuild the third row of every neighbour link, so they agree
// to rounding, not bit for bit.
template <typename R> static double shape_tolerance() { return sizeof(R) == 8 ? 1e-11 : 5e-5; }

// max |a - b| over n reals of type R (device arrays) into *out (host)
template <typename R> static int max_diff(const R *a, const R *b, int64_t n, double *part, double *out)
{
    hipLaunchKernelGGL(k_maxdiff<R>, dim3(REDUCE_BLOCKS), dim3(REDUCE_THREADS), 0, stream(), a, b, n, part);
    NRQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(REDUCE_THREADS), 0, stream(), part, REDUCE_BLOCKS, 1, 0, part + REDUCE_BLOCKS);
    NRQ_HIP(hipGetLastError());
    NRQ_HIP(hipMemcpyAsync(out, part + REDUCE_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    return NRQ_OK;
}

// Times every launch shape on scratch copies of the configuration (the
// handle's own state is not touched) and returns the fastest of those whose
// one-pass result agrees with LF_DEFAULT's (shape_tolerance); a shape that
// does not agree is reported on stderr and never chosen.
template <typename R> static int tune_leapfrog(const nrq_gauge *h, R dt, int *best)
{
    const geom g = lat().g;
    const int64_t nl = (int64_t)3 * 9 * 2 * g.vol, ne = (int64_t)3 * 8 * g.vol; // reals
    const size_t lb = (size_t)nl * sizeof(R), eb = (size_t)ne * sizeof(R);
    *best = LF_DEFAULT;
    // u0, e0: the state; ua, ea: LF_DEFAULT's pass (then timing scratch); ub, eb: the shape's pass
    void *buf[6] = {};
    const size_t sz[6] = {lb, lb, lb, eb, eb, eb};
    double *part = nullptr;
    bool ok = hipMalloc(&part, sizeof(double) * (REDUCE_BLOCKS + 1)) == hipSuccess;
    for (int k = 0; k < 6 && ok; ++k) ok = hipMalloc(&buf[k], sz[k]) == hipSuccess;
    auto release_all = [&]() {
        for (void *b : buf)
            if (b) (void)hipFree(b);
        if (part) (void)hipFree(part);
    };
    if (!ok) {
        release_all();
        (void)hipGetLastError();
        return NRQ_OK; // no room to tune: keep the default shape
    }
    cplx<R> *u0 = (cplx<R> *)buf[0], *ua = (cplx<R> *)buf[1], *ub = (cplx<R> *)buf[2];
    R *e0 = (R *)buf[3], *ea = (R *)buf[4], *eb_ = (R *)buf[5];
    NRQ_HIP(hipMemcpyAsync(u0, h->u[h->cur], lb, hipMemcpyDeviceToDevice, stream()));
    NRQ_HIP(hipMemcpyAsync(e0, h->e, eb, hipMemcpyDeviceToDevice, stream()));
    NRQ_HIP(hipMemcpyAsync(ea, e0, eb, hipMemcpyDeviceToDevice, stream()));
    launch_fused<R>(LF_DEFAULT, one<R>(u0, ua, ea), 1, g, dt);
    NRQ_HIP(hipGetLastError());
    // the scale of E: max |E - 0| after LF_DEFAULT's pass
    double emax = 0;
    NRQ_HIP(hipMemsetAsync(eb_, 0, eb, stream()));
    if (int r = max_diff<R>(ea, eb_, ne, part, &emax)) return release_all(), r;
    const double tol = shape_tolerance<R>();
    std::vector<char> valid(LF_NVARIANTS, 0);
    for (int v = 0; v < LF_NVARIANTS; ++v) {
        if (!variant_ok(v, g, (int)(8 * sizeof(R)))) continue;
        NRQ_HIP(hipMemcpyAsync(eb_, e0, eb, hipMemcpyDeviceToDevice, stream()));
        launch_fused<R>(v, one<R>(u0, ub, eb_), 1, g, dt);
        NRQ_HIP(hipGetLastError());
        double du = 0, de = 0;
        if (int r = max_diff<R>((const R *)ua, (const R *)ub, nl, part, &du)) return release_all(), r;
        if (int r = max_diff<R>(ea, eb_, ne, part, &de)) return release_all(), r;
        valid[v] = du <= tol && de <= tol * std::max(1.0, emax);
        if (!valid[v])
            fprintf(stderr,
                    "nrqcd_amd autotune: launch shape %d rejected: one pass differs from shape %d by %.3e (links), "
                    "%.3e (E)\n",
                    v, LF_DEFAULT, du, de);
    }
    hipEvent_t t0, t1;
    NRQ_HIP(hipEventCreate(&t0));
    NRQ_HIP(hipEventCreate(&t1));
    // two passes over all shapes, each shape scored by its faster pass, so
    // that one slow sample (clock ramp, another process) does not decide
    const int reps = g.vol >= (1 << 20) ? 6 : (g.vol >= (1 << 18) ? 16 : 40);
    std::vector<float> best_of(LF_NVARIANTS, 1e30f);
    for (int pass = 0; pass < 2; ++pass)
        for (int v = 0; v < LF_NVARIANTS; ++v) {
            if (!valid[v]) continue;
            launch_fused<R>(v, one<R>(u0, ua, eb_), 1, g, dt); // warm-up
            NRQ_HIP(hipEventRecord(t0, stream()));
            for (int k = 0; k < reps; ++k) launch_fused<R>(v, (k & 1) ? one<R>(ua, u0, eb_) : one<R>(u0, ua, eb_), 1, g, dt);
            NRQ_HIP(hipEventRecord(t1, stream()));
            NRQ_HIP(hipEventSynchronize(t1));
            float ms = 0;
            NRQ_HIP(hipEventElapsedTime(&ms, t0, t1));
            best_of[v] = std::min(best_of[v], ms);
            if (getenv("NRQ_AUTOTUNE_VERBOSE"))
                fprintf(stderr,
                        "nrqcd_amd autotune: precision %d V=%lld shape %d bs=%d xcd=%d tiling=%d streaming=%d two-row=%d: %.2f us/step\n",
                        (int)(8 * sizeof(R)), (long long)g.vol, v, LF_VARIANTS[v].bs, LF_VARIANTS[v].xcd, LF_VARIANTS[v].mode,
                        LF_VARIANTS[v].nt, LF_VARIANTS[v].rec, 1e3 * ms / reps);
        }
    for (int v = 0; v < LF_NVARIANTS; ++v)
        if (valid[v] && best_of[v] < best_of[*best]) *best = v;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    release_all();
    return NRQ_OK;
}

// Launch shape pinned for every leapfrog call (-1: the autotuner's choice).
// NRQ_LF_SHAPE=<n> in the environment gives the initial value (used where
// the lattice allows shape n); nrq_leapfrog_set_shape changes it.
static int g_pinned = -2; // -2: environment not read yet
static bool g_pinned_by_env = false;
static int pinned_shape()
{
    if (g_pinned == -2) {
        g_pinned = -1;
        if (const char *e = getenv("NRQ_LF_SHAPE")) {
            char *end = nullptr;
            const long v = strtol(e, &end, 10);
            if (end != e && *end == '\0' && v >= 0 && v < LF_NVARIANTS) {
                g_pinned = (int)v;
                g_pinned_by_env = true;
            }
        }
    }
    return g_pinned;
}

template <typename R> static int leapfrog_variant(const nrq_gauge *h, R dt, int *variant)
{
    if (g_autotune < 0) {
        const char *env = getenv("NRQ_AUTOTUNE");
        g_autotune = (env && env[0] == '0') ? 0 : 1;
    }
    const geom &g = lat().g;
    const int f = pinned_shape();
    if (f >= 0) {
        if (variant_ok(f, g, h->precision)) {
            *variant = f;
            return NRQ_OK;
        }
        if (!g_pinned_by_env)
            return fail(NRQ_ERR_UNSUPPORTED, "leapfrog: pinned launch shape %d does not apply to a %dx%dx%d lattice", f,
                        g.n0, g.n1, g.n2);
    }
    *variant = LF_DEFAULT;
    if (!g_autotune) return NRQ_OK;
    for (const lf_choice &c : g_lf_choice)
        if (c.precision == h->precision && c.n[0] == g.n0 && c.n[1] == g.n1 && c.n[2] == g.n2) {
            *variant = c.variant;
            return NRQ_OK;
        }
    if (int r = tune_leapfrog<R>(h, dt, variant)) return r;
    g_lf_choice.push_back({h->precision, {g.n0, g.n1, g.n2}, *variant});
    return NRQ_OK;
}

// Leapfrog of nb <= LF_MAXB configurations (same precision, current lattice)
// by n steps: one launch per pass covers all of them (blockIdx.y selects the
// configuration).  With `graph`, the launch sequence of the n steps is
// captured into a hipGraph once and replayed on later calls with the same
// buffers, parities, step count and width.
struct lf_graph {
    std::vector<const void *> key;
    int64_t n;
    double sw;
    hipGraphExec_t exec;
};
static std::vector<lf_graph> g_lf_graphs;
constexpr size_t LF_GRAPH_CACHE = 8;
// Longer sequences are launched directly: a pass takes far longer than its
// launch, and a graph of thousands of nodes costs memory and capture time.
constexpr int64_t LF_GRAPH_MAX_STEPS = 256;

// Span timer of the fused leapfrog launches (measurement only): while it is
// on, every leapfrog call that is not graph-captured records one event pair
// on the library stream around its run of fused launches, into the next of
// the reserved spans (runs past the last span are counted as skipped).
struct lf_timer {
    std::vector<hipEvent_t> ev; // start, stop per span
    int64_t used = 0, launches = 0, site_steps = 0, skipped = 0;
    bool on = false;
};
static lf_timer g_lf_timer;

// n leapfrog steps of nb configurations.  Forward steps only: with
// reu_every > 0 the links after steps reu_first, reu_first + reu_every, ...
// (< n) are projected onto SU(3) inside the fused pass that follows them.
template <typename R>
static int leapfrog_batch_t(nrq_gauge *const *hs, int nb, double sw, int64_t n, bool graph, int64_t reu_first = 0,
                            int64_t reu_every = 0)
{
    const geom g = lat().g;
    int v = 0;
    if (int r = leapfrog_variant<R>(hs[0], (R)sw, &v)) return r;
    const R dt = (R)sw;
    auto table = [&](const std::vector<int> &cur, bool swap) {
        lf_batch<R> b = {};
        for (int i = 0; i < nb; ++i) {
            b.uin[i] = (const cplx<R> *)hs[i]->u[cur[i]];
            b.uout[i] = (cplx<R> *)hs[i]->u[swap ? cur[i] ^ 1 : cur[i]];
            b.ef[i] = (R *)hs[i]->e;
        }
        return b;
    };
    std::vector<int> cur(nb);
    for (int i = 0; i < nb; ++i) cur[i] = hs[i]->cur;
    const int64_t flips = sw > 0 ? n - 1 : n;
    auto issue = [&](bool timed) {
        std::vector<int> c = cur;
        const int64_t nfused = sw > 0 ? n - 1 : n;
        lf_timer &tm = g_lf_timer;
        bool rec = false;
        if (timed && tm.on && nfused > 0) {
            rec = tm.used < (int64_t)tm.ev.size() / 2;
            if (!rec) tm.skipped += nfused;
        }
        if (sw > 0) launch_links<R>(v, table(c, false), nb, g, dt);
        if (rec) (void)hipEventRecord(tm.ev[2 * tm.used], stream());
        for (int64_t k = 0; k < nfused; ++k) {
            // forward: this pass updates E of step k + 1, then the links of step k + 2
            const int64_t st = k + 1;
            const int reu = sw > 0 && reu_every > 0 && st >= reu_first && (st - reu_first) % reu_every == 0;
            launch_fused<R>(v, table(c, true), nb, g, dt, reu);
            for (int &x : c) x ^= 1;
        }
        if (rec) {
            (void)hipEventRecord(tm.ev[2 * tm.used + 1], stream());
            ++tm.used;
            tm.launches += nfused;
            tm.site_steps += nfused * nb * g.vol;
        }
        if (sw > 0) launch_efield<R>(v, table(c, false), nb, g, dt);
    };
    if (!graph || n < 2 || n > LF_GRAPH_MAX_STEPS) {
        issue(true);
        NRQ_HIP(hipGetLastError());
    } else {
        // the captured launches carry the extents, the spacings (staple
        // coefficients), the shape and the buffers by value
        std::vector<const void *> key = {(const void *)(intptr_t)g.n0, (const void *)(intptr_t)g.n1,
                                         (const void *)(intptr_t)g.n2, (const void *)(intptr_t)v,
                                         (const void *)(intptr_t)sizeof(R), (const void *)(intptr_t)reu_first,
                                         (const void *)(intptr_t)reu_every};
        for (int k = 0; k < 4; ++k) {
            uint64_t bits;
            memcpy(&bits, &g.a[k], sizeof bits);
            key.push_back((const void *)(uintptr_t)bits);
        }
        for (int i = 0; i < nb; ++i) {
            key.push_back(hs[i]->u[cur[i]]);
            key.push_back(hs[i]->u[cur[i] ^ 1]);
            key.push_back(hs[i]->e);
        }
        hipGraphExec_t exec = nullptr;
        for (const lf_graph &c : g_lf_graphs)
            if (c.n == n && c.sw == sw && c.key == key) exec = c.exec;
        if (!exec) {
            hipGraph_t gr = nullptr;
            NRQ_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
            issue(false);
            const hipError_t le = hipGetLastError();
            const hipError_t ce = hipStreamEndCapture(stream(), &gr);
            NRQ_HIP(le);
            NRQ_HIP(ce);
            const hipError_t ie = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            NRQ_HIP(ie);
            if (g_lf_graphs.size() >= LF_GRAPH_CACHE) {
                (void)hipGraphExecDestroy(g_lf_graphs.front().exec);
                g_lf_graphs.erase(g_lf_graphs.begin());
            }
            g_lf_graphs.push_back({key, n, sw, exec});
        }
        NRQ_HIP(hipGraphLaunch(exec, stream()));
    }
    for (int i = 0; i < nb; ++i)
        if (flips & 1) hs[i]->cur ^= 1;
    return NRQ_OK;
}

template <typename R> static int leapfrog_t(nrq_gauge *h, double sw, int64_t n, int64_t reu_first = 0, int64_t reu_every = 0)
{
    nrq_gauge *hs[1] = {h};
    return leapfrog_batch_t<R>(hs, 1, sw, n, false, reu_first, reu_every);
}

// Drops the captured leapfrog graphs (their launches carry the lattice by
// value); called when the lattice module is finalised.
void gauge_forget_lattice()
{
    for (lf_graph &c : g_lf_graphs) (void)hipGraphExecDestroy(c.exec);
    g_lf_graphs.clear();
}

int leapfrog_choice(const nrq_gauge *h, int *bs, int *xcd, int *mode, int *nt, int *rec)
{
    int v = 0;
    if (int r = h->precision == 64 ? leapfrog_variant<double>(h, 1.0, &v) : leapfrog_variant<float>(h, 1.0f, &v))
        return r;
    *bs = LF_VARIANTS[v].bs;
    *xcd = LF_VARIANTS[v].xcd;
    *mode = LF_VARIANTS[v].mode;
    *nt = LF_VARIANTS[v].nt;
    *rec = LF_VARIANTS[v].rec;
    return NRQ_OK;
}

} // namespace nrq

using namespace nrq;

#define NEED_LATTICE()                                                                  \
    do {                                                                                \
        if (!lat().initialised) return fail(NRQ_ERR_STATE, "lattice: not initialised"); \
    } while (0)

#define NEED_GAUGE(h)                                                                          \
    do {                                                                                       \
        if (!(h)) return fail(NRQ_ERR_ARG, "gauge: null handle");                              \
        if (!(h)->u[0]) return fail(NRQ_ERR_STATE, "gauge: not allocated");                    \
        if ((h)->vol != lat().g.vol) return fail(NRQ_ERR_STATE, "gauge: lattice changed since allocation"); \
    } while (0)

extern "C" int nrq_gauge_create(int precision, nrq_gauge **out)
{
    NEED_LATTICE();
    if (!out) return fail(NRQ_ERR_ARG, "gauge: null output handle");
    if (precision != 64 && precision != 32) return fail(NRQ_ERR_ARG, "gauge: precision must be 64 or 32");
    if (int r = ensure_device()) return r;
    nrq_gauge *h = new nrq_gauge();
    if (int r = allocate(h, precision)) {
        delete h;
        return r;
    }
    // defined contents: cold configuration
    *out = h;
    return nrq_gauge_cold(h);
}

extern "C" int nrq_gauge_destroy(nrq_gauge *h)
{
    if (!h) return NRQ_OK;
    if (stream()) (void)hipStreamSynchronize(stream());
    release(h);
    delete h;
    return NRQ_OK;
}

extern "C" int nrq_gauge_copy(nrq_gauge *dst, const nrq_gauge *src)
{
    NEED_GAUGE(src);
    if (!dst) return fail(NRQ_ERR_ARG, "gauge: null destination");
    if (dst == src) return NRQ_OK;
    if (!dst->u[0] || dst->precision != src->precision || dst->vol != src->vol)
        if (int r = allocate(dst, src->precision)) return r;
    NRQ_HIP(hipMemcpyAsync(dst->u[0], src->u[src->cur], link_bytes(src), hipMemcpyDeviceToDevice, stream()));
    NRQ_HIP(hipMemcpyAsync(dst->e, src->e, efield_bytes(src), hipMemcpyDeviceToDevice, stream()));
    dst->cur = 0;
    dst->steps_since_reunit = src->steps_since_reunit;
    dst->reunit_interval = src->reunit_interval;
    return NRQ_OK;
}

extern "C" int nrq_gauge_cold(nrq_gauge *h)
{
    NEED_GAUGE(h);
    const int64_t v = h->vol;
    h->cur = 0;
    if (h->precision == 64)
        hipLaunchKernelGGL(k_cold<double>, dim3((unsigned)((v + 255) / 256)), dim3(256), 0, stream(),
                           (cplx<double> *)h->u[0], (double *)h->e, v);
    else
        hipLaunchKernelGGL(k_cold<float>, dim3((unsigned)((v + 255) / 256)), dim3(256), 0, stream(),
                           (cplx<float> *)h->u[0], (float *)h->e, v);
    NRQ_HIP(hipGetLastError());
    return NRQ_OK;
}

extern "C" int nrq_gauge_upload(nrq_gauge *h, const double *links, const double *efield)
{
    NEED_GAUGE(h);
    const int64_t v = h->vol;
    const bool dp = h->precision == 64;
    if (links) {
        std::vector<char> buf(link_bytes(h));
        for (int64_t s = 0; s < v; ++s)
            for (int i = 0; i < 3; ++i)
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) {
                        const double *src = links + 2 * (r + 3 * c + 9 * i + 27 * s);
                        const size_t at = ((size_t)i * 9 + 3 * r + c) * (size_t)v + (size_t)s;
                        if (dp) {
                            ((double *)buf.data())[2 * at] = src[0];
                            ((double *)buf.data())[2 * at + 1] = src[1];
                        } else {
                            ((float *)buf.data())[2 * at] = (float)src[0];
                            ((float *)buf.data())[2 * at + 1] = (float)src[1];
                        }
                    }
        NRQ_HIP(hipMemcpyAsync(h->u[h->cur], buf.data(), buf.size(), hipMemcpyHostToDevice, stream()));
        NRQ_HIP(hipStreamSynchronize(stream()));
    }
    if (efield) {
        std::vector<char> buf(efield_bytes(h));
        for (int64_t s = 0; s < v; ++s)
            for (int i = 0; i < 3; ++i)
                for (int a = 0; a < 8; ++a) {
                    const double x = efield[a + 8 * i + 24 * s];
                    const size_t at = ((size_t)i * 8 + a) * (size_t)v + (size_t)s;
                    if (dp) ((double *)buf.data())[at] = x;
                    else ((float *)buf.data())[at] = (float)x;
                }
        NRQ_HIP(hipMemcpyAsync(h->e, buf.data(), buf.size(), hipMemcpyHostToDevice, stream()));
        NRQ_HIP(hipStreamSynchronize(stream()));
    }
    return NRQ_OK;
}

extern "C" int nrq_gauge_download(const nrq_gauge *h, double *links, double *efield)
{
    NEED_GAUGE(h);
    const int64_t v = h->vol;
    const bool dp = h->precision == 64;
    if (links) {
        std::vector<char> buf(link_bytes(h));
        NRQ_HIP(hipMemcpyAsync(buf.data(), h->u[h->cur], buf.size(), hipMemcpyDeviceToHost, stream()));
        NRQ_HIP(hipStreamSynchronize(stream()));
        for (int64_t s = 0; s < v; ++s)
            for (int i = 0; i < 3; ++i)
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) {
                        double *dst = links + 2 * (r + 3 * c + 9 * i + 27 * s);
                        const size_t at = ((size_t)i * 9 + 3 * r + c) * (size_t)v + (size_t)s;
                        if (dp) {
                            dst[0] = ((double *)buf.data())[2 * at];
                            dst[1] = ((double *)buf.data())[2 * at + 1];
                        } else {
                            dst[0] = ((float *)buf.data())[2 * at];
                            dst[1] = ((float *)buf.data())[2 * at + 1];
                        }
                    }
    }
    if (efield) {
        std::vector<char> buf(efield_bytes(h));
        NRQ_HIP(hipMemcpyAsync(buf.data(), h->e, buf.size(), hipMemcpyDeviceToHost, stream()));
        NRQ_HIP(hipStreamSynchronize(stream()));
        for (int64_t s = 0; s < v; ++s)
            for (int i = 0; i < 3; ++i)
                for (int a = 0; a < 8; ++a) {
                    const size_t at = ((size_t)i * 8 + a) * (size_t)v + (size_t)s;
                    efield[a + 8 * i + 24 * s] = dp ? ((double *)buf.data())[at] : ((float *)buf.data())[at];
                }
    }
    return NRQ_OK;
}

extern "C" int nrq_gauge_get_link(const nrq_gauge *h, int dir, int64_t idx, double link[18])
{
    NEED_GAUGE(h);
    if (!link) return fail(NRQ_ERR_ARG, "null argument");
    if (dir < -3 || dir > 3) return fail(NRQ_ERR_ARG, "direction %d out of range", dir);
    if (idx < 1 || idx > h->vol) return fail(NRQ_ERR_ARG, "lattice index out of range");
    int64_t s = idx - 1;
    if (dir < 0) {
        int64_t nb;
        if (int r = nrq_lattice_neib(dir, idx, &nb)) return r;
        s = nb - 1;
    }
    double *dres = (double *)h->partials + NOBS * REDUCE_BLOCKS;
    if (h->precision == 64)
        hipLaunchKernelGGL(k_get_link<double>, dim3(1), dim3(64), 0, stream(), (const cplx<double> *)h->u[h->cur], h->vol, dir, s, dres);
    else
        hipLaunchKernelGGL(k_get_link<float>, dim3(1), dim3(64), 0, stream(), (const cplx<float> *)h->u[h->cur], h->vol, dir, s, dres);
    NRQ_HIP(hipGetLastError());
    NRQ_HIP(hipMemcpyAsync(h->result, dres, 18 * sizeof(double), hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    memcpy(link, h->result, 18 * sizeof(double));
    return NRQ_OK;
}

extern "C" int nrq_gauge_get_efield(const nrq_gauge *h, int dir, int64_t idx, double e[8])
{
    NEED_GAUGE(h);
    if (!e) return fail(NRQ_ERR_ARG, "null argument");
    if (dir < 1 || dir > 3) return fail(NRQ_ERR_ARG, "direction %d out of range", dir);
    if (idx < 1 || idx > h->vol) return fail(NRQ_ERR_ARG, "lattice index out of range");
    const size_t es = h->precision == 64 ? 8 : 4;
    char tmp[8][8];
    for (int a = 0; a < 8; ++a)
        NRQ_HIP(hipMemcpyAsync(tmp[a], (char *)h->e + (((size_t)(dir - 1) * 8 + a) * (size_t)h->vol + (size_t)(idx - 1)) * es, es,
                               hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    for (int a = 0; a < 8; ++a) e[a] = es == 8 ? *(double *)tmp[a] : (double)*(float *)tmp[a];
    return NRQ_OK;
}

extern "C" int nrq_gauge_leapfrog(nrq_gauge *h, double stepwidth, int64_t nsteps)
{
    NEED_GAUGE(h);
    if (nsteps < 0 || !std::isfinite(stepwidth)) return fail(NRQ_ERR_ARG, "leapfrog: bad step count or width");
    if (stepwidth > 0 && h->reunit_interval > 0 && nsteps > 0) {
        // forward steps: the reunitarisation points inside the call are done by
        // the fused passes that follow them (the links after step k are
        // projected between the E update of step k and the link update of step
        // k + 1, as a separate pass there would); one at the call's end by a
        // separate pass.  Points after 1, 1 + interval, ... steps when the count
        // has already reached the interval.
        const int64_t iv = h->reunit_interval;
        const int64_t first = std::max<int64_t>(1, iv - h->steps_since_reunit);
        if (int r = h->precision == 64 ? leapfrog_t<double>(h, stepwidth, nsteps, first, iv)
                                       : leapfrog_t<float>(h, stepwidth, nsteps, first, iv))
            return r;
        if (nsteps < first) {
            h->steps_since_reunit += nsteps;
            return NRQ_OK;
        }
        const int64_t last = first + (nsteps - first) / iv * iv; // last point <= nsteps
        h->steps_since_reunit = nsteps - last;
        if (last == nsteps) return nrq_gauge_reunitarise(h);
        return NRQ_OK;
    }
    while (nsteps > 0) {
        // with a reunitarisation interval, stop at each multiple of it
        int64_t n = nsteps;
        if (h->reunit_interval > 0) n = std::min(n, std::max<int64_t>(1, h->reunit_interval - h->steps_since_reunit));
        if (int r = h->precision == 64 ? leapfrog_t<double>(h, stepwidth, n) : leapfrog_t<float>(h, stepwidth, n)) return r;
        h->steps_since_reunit += n;
        nsteps -= n;
        if (h->reunit_interval > 0 && h->steps_since_reunit >= h->reunit_interval)
            if (int r = nrq_gauge_reunitarise(h)) return r;
    }
    return NRQ_OK;
}

extern "C" int nrq_leapfrog_timer_start(int64_t max_spans)
{
    if (max_spans < 1 || max_spans > 1000000) return fail(NRQ_ERR_ARG, "leapfrog timer: bad span count %lld", (long long)max_spans);
    if (int r = ensure_device()) return r;
    lf_timer &tm = g_lf_timer;
    while ((int64_t)tm.ev.size() < 2 * max_spans) {
        hipEvent_t e = nullptr;
        NRQ_HIP(hipEventCreate(&e));
        tm.ev.push_back(e);
    }
    tm.used = tm.launches = tm.site_steps = tm.skipped = 0;
    tm.on = true;
    return NRQ_OK;
}

extern "C" int nrq_leapfrog_timer_stop(double *fused_ms, int64_t *launches, int64_t *site_steps, int64_t *skipped)
{
    lf_timer &tm = g_lf_timer;
    if (!tm.on) return fail(NRQ_ERR_STATE, "leapfrog timer: not started");
    tm.on = false;
    NRQ_HIP(hipStreamSynchronize(stream()));
    double ms = 0;
    for (int64_t k = 0; k < tm.used; ++k) {
        float t = 0;
        NRQ_HIP(hipEventElapsedTime(&t, tm.ev[2 * k], tm.ev[2 * k + 1]));
        ms += t;
    }
    if (fused_ms) *fused_ms = ms;
    if (launches) *launches = tm.launches;
    if (site_steps) *site_steps = tm.site_steps;
    if (skipped) *skipped = tm.skipped;
    return NRQ_OK;
}

extern "C" int nrq_gauge_leapfrog_batch(nrq_gauge *const *h, int n, double stepwidth, int64_t nsteps)
{
    if (!h || n < 1) return fail(NRQ_ERR_ARG, "leapfrog batch: no configurations");
    if (nsteps < 0 || !std::isfinite(stepwidth)) return fail(NRQ_ERR_ARG, "leapfrog: bad step count or width");
    for (int i = 0; i < n; ++i) {
        NEED_GAUGE(h[i]);
        if (h[i]->precision != h[0]->precision) return fail(NRQ_ERR_ARG, "leapfrog batch: mixed storage precisions");
        for (int j = 0; j < i; ++j)
            if (h[j] == h[i]) return fail(NRQ_ERR_ARG, "leapfrog batch: configuration %d listed twice", i);
    }
    const char *env = getenv("NRQ_GRAPH");
    const bool graph = !(env && env[0] == '0');
    bool uniform = stepwidth > 0 && nsteps > 0 && h[0]->reunit_interval > 0;
    for (int i = 1; i < n && uniform; ++i)
        uniform = h[i]->reunit_interval == h[0]->reunit_interval && h[i]->steps_since_reunit == h[0]->steps_since_reunit;
    if (uniform) {
        // every member at the same point of the same interval (forward steps):
        // the points inside the call are done by the fused passes, as in
        // nrq_gauge_leapfrog, one at the call's end by a separate pass
        const int64_t iv = h[0]->reunit_interval;
        const int64_t first = std::max<int64_t>(1, iv - h[0]->steps_since_reunit);
        for (int i0 = 0; i0 < n; i0 += LF_MAXB) {
            const int nb = std::min(LF_MAXB, n - i0);
            if (int r = h[0]->precision == 64 ? leapfrog_batch_t<double>(h + i0, nb, stepwidth, nsteps, graph, first, iv)
                                              : leapfrog_batch_t<float>(h + i0, nb, stepwidth, nsteps, graph, first, iv))
                return r;
        }
        for (int i = 0; i < n; ++i) {
            if (nsteps < first) {
                h[i]->steps_since_reunit += nsteps;
                continue;
            }
            const int64_t last = first + (nsteps - first) / iv * iv;
            h[i]->steps_since_reunit = nsteps - last;
            if (last == nsteps)
                if (int r = nrq_gauge_reunitarise(h[i])) return r;
        }
        return NRQ_OK;
    }
    while (nsteps > 0) {
        // stop at the next reunitarisation point of any configuration
        int64_t m = nsteps;
        for (int i = 0; i < n; ++i)
            if (h[i]->reunit_interval > 0)
                m = std::min(m, std::max<int64_t>(1, h[i]->reunit_interval - h[i]->steps_since_reunit));
        for (int i0 = 0; i0 < n; i0 += LF_MAXB) {
            const int nb = std::min(LF_MAXB, n - i0);
            if (int r = h[0]->precision == 64 ? leapfrog_batch_t<double>(h + i0, nb, stepwidth, m, graph)
                                              : leapfrog_batch_t<float>(h + i0, nb, stepwidth, m, graph))
                return r;
        }
        nsteps -= m;
        for (int i = 0; i < n; ++i) {
            h[i]->steps_since_reunit += m;
            if (h[i]->reunit_interval > 0 && h[i]->steps_since_reunit >= h[i]->reunit_interval)
                if (int r = nrq_gauge_reunitarise(h[i])) return r;
        }
    }
    return NRQ_OK;
}

extern "C" int nrq_gauge_set_reunitarise_interval(nrq_gauge *h, int64_t interval)
{
    NEED_GAUGE(h);
    if (interval < 0) return fail(NRQ_ERR_ARG, "reunitarisation interval %lld < 0", (long long)interval);
    h->reunit_interval = interval;
    return NRQ_OK;
}

extern "C" int nrq_gauge_observables(const nrq_gauge *h, double out[4])
{
    NEED_GAUGE(h);
    if (!out) return fail(NRQ_ERR_ARG, "null argument");
    double o[NOBS];
    if (int r = gauge_obs(h, o)) return r;
    obs_row(o, h->vol, out);
    return NRQ_OK;
}

extern "C" int nrq_gauge_log_observables(nrq_gauge *h)
{
    NEED_GAUGE(h);
    if (h->log_n == h->log_cap) {
        // grow by doubling (at least 64 rows), keeping the rows held
        const int64_t cap = h->log_cap ? 2 * h->log_cap : 64;
        double *d = nullptr;
        if (hipMalloc((void **)&d, sizeof(double) * NOBS * (size_t)cap) != hipSuccess) {
            (void)hipGetLastError();
            return fail(NRQ_ERR_ALLOC, "gauge: observables log of %lld rows", (long long)cap);
        }
        if (h->log_n)
            NRQ_HIP(hipMemcpyAsync(d, h->dlog, sizeof(double) * NOBS * (size_t)h->log_n, hipMemcpyDeviceToDevice, stream()));
        NRQ_HIP(hipStreamSynchronize(stream()));
        if (h->dlog) (void)hipFree(h->dlog);
        h->dlog = d;
        h->log_cap = cap;
    }
    if (int r = gauge_obs_enqueue(h, h->dlog + NOBS * h->log_n)) return r;
    ++h->log_n;
    return NRQ_OK;
}

extern "C" int nrq_gauge_log_read(nrq_gauge *h, double *out, int64_t max_rows, int64_t *rows)
{
    NEED_GAUGE(h);
    if (!rows || (max_rows > 0 && !out) || max_rows < 0) return fail(NRQ_ERR_ARG, "null or negative argument");
    const int64_t n = std::min(h->log_n, max_rows);
    *rows = n;
    if (n == 0) return NRQ_OK;
    std::vector<double> raw((size_t)NOBS * (size_t)n);
    NRQ_HIP(hipMemcpyAsync(raw.data(), h->dlog, sizeof(double) * raw.size(), hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    for (int64_t r = 0; r < n; ++r) obs_row(raw.data() + NOBS * r, h->vol, out + 4 * r);
    const int64_t left = h->log_n - n;
    if (left > 0) {
        // the rows not returned move to the front (non-overlapping when left <= n; else in n-row pieces)
        for (int64_t k = 0; k < left; k += n) {
            const int64_t m = std::min(n, left - k);
            NRQ_HIP(hipMemcpyAsync(h->dlog + NOBS * k, h->dlog + NOBS * (n + k), sizeof(double) * NOBS * (size_t)m,
                                   hipMemcpyDeviceToDevice, stream()));
        }
    }
    h->log_n = left;
    return NRQ_OK;
}

extern "C" int nrq_gauge_energy(const nrq_gauge *h, double *v)
{
    double o[4];
    if (!v) return fail(NRQ_ERR_ARG, "null argument");
    if (int r = nrq_gauge_observables(h, o)) return r;
    *v = o[0];
    return NRQ_OK;
}

extern "C" int nrq_gauge_gauss_ref(const nrq_gauge *h, double *v)
{
    double o[4];
    if (!v) return fail(NRQ_ERR_ARG, "null argument");
    if (int r = nrq_gauge_observables(h, o)) return r;
    *v = o[1];
    return NRQ_OK;
}

extern "C" int nrq_gauge_gauss_max(const nrq_gauge *h, double *v)
{
    double o[4];
    if (!v) return fail(NRQ_ERR_ARG, "null argument");
    if (int r = nrq_gauge_observables(h, o)) return r;
    *v = o[2];
    return NRQ_OK;
}

extern "C" int nrq_gauge_mean_plaquette(const nrq_gauge *h, double *v)
{
    double o[4];
    if (!v) return fail(NRQ_ERR_ARG, "null argument");
    if (int r = nrq_gauge_observables(h, o)) return r;
    *v = o[3];
    return NRQ_OK;
}

extern "C" int nrq_gauge_plaquette(const nrq_gauge *h, int i, int j, int64_t idx, double P[18])
{
    NEED_GAUGE(h);
    if (!P) return fail(NRQ_ERR_ARG, "null argument");
    if (i < 1 || i > 3 || j < 1 || j > 3) return fail(NRQ_ERR_ARG, "plaquette: spatial directions 1..3 only");
    if (idx < 1 || idx > h->vol) return fail(NRQ_ERR_ARG, "lattice index out of range");
    double *dres = (double *)h->partials + NOBS * REDUCE_BLOCKS;
    const geom g = lat().g;
    if (h->precision == 64)
        hipLaunchKernelGGL(k_plaq<double>, dim3(1), dim3(64), 0, stream(), (const cplx<double> *)h->u[h->cur], g, i - 1, j - 1, idx - 1, dres);
    else
        hipLaunchKernelGGL(k_plaq<float>, dim3(1), dim3(64), 0, stream(), (const cplx<float> *)h->u[h->cur], g, i - 1, j - 1, idx - 1, dres);
    NRQ_HIP(hipGetLastError());
    NRQ_HIP(hipMemcpyAsync(h->result, dres, 18 * sizeof(double), hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    memcpy(P, h->result, 18 * sizeof(double));
    return NRQ_OK;
}

extern "C" int nrq_gauge_reunitarise(nrq_gauge *h)
{
    NEED_GAUGE(h);
    const dim3 grd((unsigned)((h->vol + 255) / 256), 3);
    if (h->precision == 64)
        hipLaunchKernelGGL(k_reunit<double>, grd, dim3(256), 0, stream(), (cplx<double> *)h->u[h->cur], h->vol);
    else
        hipLaunchKernelGGL(k_reunit<float>, grd, dim3(256), 0, stream(), (cplx<float> *)h->u[h->cur], h->vol);
    NRQ_HIP(hipGetLastError());
    h->steps_since_reunit = 0;
    return NRQ_OK;
}

// Accessors used by the other translation units.
namespace nrq {
const void *gauge_links(const nrq_gauge *h) { return h->u[h->cur]; }
void *gauge_links_mut(nrq_gauge *h) { return h->u[h->cur]; }
void *gauge_efield(nrq_gauge *h) { return h->e; }
}

extern "C" int nrq_gauge_launch_shape(const nrq_gauge *h, int *block_sites, int *xcd_remap, int *tiling, int *streaming,
                                      int *two_row)
{
    NEED_GAUGE(h);
    if (!block_sites || !xcd_remap || !tiling || !streaming || !two_row) return fail(NRQ_ERR_ARG, "null argument");
    return leapfrog_choice(h, block_sites, xcd_remap, tiling, streaming, two_row);
}

extern "C" int nrq_leapfrog_shape_count(int *n)
{
    if (!n) return fail(NRQ_ERR_ARG, "null argument");
    *n = LF_NVARIANTS;
    return NRQ_OK;
}

extern "C" int nrq_leapfrog_shape_info(int k, int *block_sites, int *xcd_remap, int *tiling, int *streaming, int *two_row,
                                       int *applies)
{
    if (k < 0 || k >= LF_NVARIANTS) return fail(NRQ_ERR_ARG, "launch shape %d out of range", k);
    if (!block_sites || !xcd_remap || !tiling || !streaming || !two_row || !applies) return fail(NRQ_ERR_ARG, "null argument");
    const lf_variant &v = LF_VARIANTS[k];
    *block_sites = v.bs;
    *xcd_remap = v.xcd;
    *tiling = v.mode;
    *streaming = v.nt;
    *two_row = v.rec;
    *applies = lat().initialised ? (variant_ok(k, lat().g, 64) ? 1 : 0) | (variant_ok(k, lat().g, 32) ? 2 : 0) : 0;
    return NRQ_OK;
}

extern "C" int nrq_leapfrog_set_shape(int k)
{
    if (k < -1 || k >= LF_NVARIANTS) return fail(NRQ_ERR_ARG, "launch shape %d out of range", k);
    (void)pinned_shape(); // read the environment once, so that it cannot override this call later
    g_pinned = k;
    g_pinned_by_env = false;
    return NRQ_OK;
}

extern "C" int nrq_leapfrog_get_shape(int *k)
{
    if (!k) return fail(NRQ_ERR_ARG, "null argument");
    *k = pinned_shape();
    return NRQ_OK;
}
