// nrq_su3.h -- device-side complex and SU(3) arithmetic for gfx950.
//
// Matrices live in registers as 9 complex numbers, row-major (m[3*r + c]).
// Everything is templated on the storage/compute real type R (double or
// float).  The generator convention is the reference's (su3.f90:97-131): the
// Fortran reshape fills column-major, so T^2, T^5 and T^7 are the complex
// conjugates of the textbook Gell-Mann matrices.  Written out, with
// s = 1/(2 sqrt 3):
//   alpha^a T^a = [[ a3/2 + a8 s,   (a1 + i a2)/2,  (a4 + i a5)/2 ],
//                  [ (a1 - i a2)/2, -a3/2 + a8 s,   (a6 + i a7)/2 ],
//                  [ (a4 - i a5)/2, (a6 - i a7)/2,  -2 a8 s       ]]
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nrq {

template <typename R> struct cplx {
    R re, im;
};

template <typename R> __device__ __forceinline__ cplx<R> mk(R a, R b) { return {a, b}; }
template <typename R> __device__ __forceinline__ cplx<R> operator+(cplx<R> a, cplx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R> __device__ __forceinline__ cplx<R> operator-(cplx<R> a, cplx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R> __device__ __forceinline__ cplx<R> operator*(cplx<R> a, cplx<R> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R> __device__ __forceinline__ cplx<R> operator*(R s, cplx<R> a) { return {s * a.re, s * a.im}; }
template <typename R> __device__ __forceinline__ cplx<R> conjg(cplx<R> a) { return {a.re, -a.im}; }
// a * conj(b)
template <typename R> __device__ __forceinline__ cplx<R> mulc(cplx<R> a, cplx<R> b)
{
    return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
// conj(a) * b
template <typename R> __device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b)
{
    return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}

// Streaming (nontemporal) access for data touched once per pass, so that it
// does not displace reused data from the caches.
template <typename R> struct vec2_of;
template <> struct vec2_of<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};
template <> struct vec2_of<float> {
    typedef float type __attribute__((ext_vector_type(2)));
};
template <typename R> __device__ __forceinline__ void store_nt(cplx<R> *p, cplx<R> v)
{
    typedef typename vec2_of<R>::type V;
    V w;
    w.x = v.re;
    w.y = v.im;
    __builtin_nontemporal_store(w, (V *)p);
}
template <typename R> __device__ __forceinline__ cplx<R> load_nt(const cplx<R> *p)
{
    typedef typename vec2_of<R>::type V;
    const V w = __builtin_nontemporal_load((const V *)p);
    return {w.x, w.y};
}

// Element i of an array whose base p is the same for the whole wavefront:
// the scalar base plus a 32-bit byte offset (i < 3 V <= 2^27.6 and 16-byte
// elements keep the offset below 2^32), so no 64-bit address arithmetic
// runs in the vector registers.
template <typename V> __device__ __forceinline__ const V *at32(const V *p, uint32_t i)
{
    return (const V *)((const char *)p + i * (uint32_t)sizeof(V));
}
template <typename V> __device__ __forceinline__ V *at32(V *p, uint32_t i)
{
    return (V *)((char *)p + i * (uint32_t)sizeof(V));
}

template <typename R> struct mat3 {
    cplx<R> m[9];
};

template <typename R> __device__ __forceinline__ void set_zero(mat3<R> &a)
{
#pragma unroll
    for (int k = 0; k < 9; ++k) a.m[k] = {R(0), R(0)};
}

// Complex multiply-accumulate as four FMAs (the plain s + a * b form
// cannot be contracted without reassociation and costs six instructions).
__device__ __forceinline__ double fmx(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmx(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// s += a b
template <typename R> __device__ __forceinline__ void mac(cplx<R> &s, cplx<R> a, cplx<R> b)
{
    s.re = fmx(a.re, b.re, s.re);
    s.re = fmx(-a.im, b.im, s.re);
    s.im = fmx(a.re, b.im, s.im);
    s.im = fmx(a.im, b.re, s.im);
}
// s += conj(a) b
template <typename R> __device__ __forceinline__ void mac_cn(cplx<R> &s, cplx<R> a, cplx<R> b)
{
    s.re = fmx(a.re, b.re, s.re);
    s.re = fmx(a.im, b.im, s.re);
    s.im = fmx(a.re, b.im, s.im);
    s.im = fmx(-a.im, b.re, s.im);
}
// s += a conj(b)
template <typename R> __device__ __forceinline__ void mac_nc(cplx<R> &s, cplx<R> a, cplx<R> b)
{
    s.re = fmx(a.re, b.re, s.re);
    s.re = fmx(a.im, b.im, s.re);
    s.im = fmx(a.im, b.re, s.im);
    s.im = fmx(-a.re, b.im, s.im);
}
// s += conj(a) conj(b)
template <typename R> __device__ __forceinline__ void mac_cc(cplx<R> &s, cplx<R> a, cplx<R> b)
{
    s.re = fmx(a.re, b.re, s.re);
    s.re = fmx(-a.im, b.im, s.re);
    s.im = fmx(-a.re, b.im, s.im);
    s.im = fmx(-a.im, b.re, s.im);
}

// C = A * B
template <typename R> __device__ __forceinline__ mat3<R> mul(const mat3<R> &a, const mat3<R> &b)
{
    mat3<R> c;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            cplx<R> s = {R(0), R(0)};
            mac(s, a.m[3 * r], b.m[k]);
            mac(s, a.m[3 * r + 1], b.m[3 + k]);
            mac(s, a.m[3 * r + 2], b.m[6 + k]);
            c.m[3 * r + k] = s;
        }
    return c;
}

// C = A * B^dagger
template <typename R> __device__ __forceinline__ mat3<R> mul_nd(const mat3<R> &a, const mat3<R> &b)
{
    mat3<R> c;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            cplx<R> s = {R(0), R(0)};
            mac_nc(s, a.m[3 * r], b.m[3 * k]);
            mac_nc(s, a.m[3 * r + 1], b.m[3 * k + 1]);
            mac_nc(s, a.m[3 * r + 2], b.m[3 * k + 2]);
            c.m[3 * r + k] = s;
        }
    return c;
}

// C = A^dagger * B
template <typename R> __device__ __forceinline__ mat3<R> mul_dn(const mat3<R> &a, const mat3<R> &b)
{
    mat3<R> c;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            cplx<R> s = {R(0), R(0)};
            mac_cn(s, a.m[r], b.m[k]);
            mac_cn(s, a.m[3 + r], b.m[3 + k]);
            mac_cn(s, a.m[6 + r], b.m[6 + k]);
            c.m[3 * r + k] = s;
        }
    return c;
}

// C = A^dagger * B^dagger = (B A)^dagger
template <typename R> __device__ __forceinline__ mat3<R> mul_dd(const mat3<R> &a, const mat3<R> &b)
{
    mat3<R> c;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            // sum_j conj(a[j][r]) conj(b[k][j])
            cplx<R> s = {R(0), R(0)};
            mac_cc(s, a.m[r], b.m[3 * k]);
            mac_cc(s, a.m[3 + r], b.m[3 * k + 1]);
            mac_cc(s, a.m[6 + r], b.m[3 * k + 2]);
            c.m[3 * r + k] = s;
        }
    return c;
}

// Rows 0 and 1 of C = A B^dagger (reads rows 0 and 1 of A; row 2 of C is
// left as it was).
template <typename R> __device__ __forceinline__ void mul_nd_top(const mat3<R> &a, const mat3<R> &b, mat3<R> &c)
{
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            cplx<R> s = {R(0), R(0)};
            mac_nc(s, a.m[3 * r], b.m[3 * k]);
            mac_nc(s, a.m[3 * r + 1], b.m[3 * k + 1]);
            mac_nc(s, a.m[3 * r + 2], b.m[3 * k + 2]);
            c.m[3 * r + k] = s;
        }
}

// Rows 0 and 1 of C = A B (reads rows 0 and 1 of A).
template <typename R> __device__ __forceinline__ void mul_top(const mat3<R> &a, const mat3<R> &b, mat3<R> &c)
{
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            cplx<R> s = {R(0), R(0)};
            mac(s, a.m[3 * r], b.m[k]);
            mac(s, a.m[3 * r + 1], b.m[3 + k]);
            mac(s, a.m[3 * r + 2], b.m[6 + k]);
            c.m[3 * r + k] = s;
        }
}

// Rows 0 and 1 of C = A^dagger B^dagger (reads columns 0 and 1 of A).
template <typename R> __device__ __forceinline__ void mul_dd_top(const mat3<R> &a, const mat3<R> &b, mat3<R> &c)
{
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            cplx<R> s = {R(0), R(0)};
            mac_cc(s, a.m[r], b.m[3 * k]);
            mac_cc(s, a.m[3 + r], b.m[3 * k + 1]);
            mac_cc(s, a.m[6 + r], b.m[3 * k + 2]);
            c.m[3 * r + k] = s;
        }
}

// Row 2 of an SU(3) matrix from rows 0 and 1: row2_c = conj(row0_{c+1}
// row1_{c+2} - row0_{c+2} row1_{c+1}) (unitarity with unit determinant).
template <typename R> __device__ __forceinline__ void su3_row3(mat3<R> &m)
{
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        cplx<R> t = {R(0), R(0)};
        mac_cc(t, m.m[c1], m.m[3 + c2]);
        const cplx<R> a = {-m.m[c2].re, -m.m[c2].im};
        mac_cc(t, a, m.m[3 + c1]);
        m.m[6 + c] = t;
    }
}

// S^U + S^D of one plane of the staple sum, with S^U = A B^+ C^+ and
// S^D = D^+ E^+ F.  Both are products of SU(3) matrices, hence SU(3): only
// their first two rows are multiplied out and row 3 is rebuilt as the
// conjugated cross product of rows 1 and 2, 48 fewer multiply-adds per
// staple.  The difference from the full product is of the order of the links'
// unitarity error (FP64 rounding for FP64 links, FP32 rounding for FP32).
template <typename R>
__device__ __forceinline__ mat3<R> staple_pair(const mat3<R> &a, const mat3<R> &b, const mat3<R> &c, const mat3<R> &d,
                                              const mat3<R> &e, const mat3<R> &f)
{
    mat3<R> su, sd, t;
    mul_nd_top(a, b, t);
    mul_nd_top(t, c, su);
    su3_row3(su);
    mul_dd_top(d, e, t);
    mul_top(t, f, sd);
    su3_row3(sd);
#pragma unroll
    for (int k = 0; k < 9; ++k) su.m[k] = su.m[k] + sd.m[k];
    return su;
}

template <typename R> __device__ __forceinline__ void add_to(mat3<R> &a, const mat3<R> &b)
{
#pragma unroll
    for (int k = 0; k < 9; ++k) a.m[k] = a.m[k] + b.m[k];
}

template <typename R> __device__ __forceinline__ cplx<R> trace(const mat3<R> &a) { return a.m[0] + a.m[4] + a.m[8]; }

// Hermitian algebra matrix alpha^a T^a (su3.f90:287-301 with the generator
// table written out in the header comment).
template <typename R> __device__ __forceinline__ mat3<R> algebra_matrix(const R *al)
{
    const R h = R(0.5);
    const R s = R(0.28867513459481288225457439025098); // 1/(2 sqrt 3)
    mat3<R> q;
    q.m[0] = {h * al[2] + s * al[7], R(0)};
    q.m[4] = {-h * al[2] + s * al[7], R(0)};
    q.m[8] = {R(-2) * s * al[7], R(0)};
    q.m[1] = {h * al[0], h * al[1]};
    q.m[3] = {h * al[0], -h * al[1]};
    q.m[2] = {h * al[3], h * al[4]};
    q.m[6] = {h * al[3], -h * al[4]};
    q.m[5] = {h * al[5], h * al[6]};
    q.m[7] = {h * al[5], -h * al[6]};
    return q;
}

// out[a] = 2 Im Tr(T^a W) for a = 1..8 (su3.f90:369-381 with the generator
// table written out in the header comment).
template <typename R> __device__ __forceinline__ void two_im_trace_gen(const mat3<R> &w, R *out)
{
    const R s3 = R(0.57735026918962576450914878050196); // 1/sqrt 3
    out[0] = w.m[1].im + w.m[3].im;
    out[1] = w.m[3].re - w.m[1].re;
    out[2] = w.m[0].im - w.m[4].im;
    out[3] = w.m[2].im + w.m[6].im;
    out[4] = w.m[6].re - w.m[2].re;
    out[5] = w.m[5].im + w.m[7].im;
    out[6] = w.m[7].re - w.m[5].re;
    out[7] = s3 * (w.m[0].im + w.m[4].im - R(2) * w.m[8].im);
}

// Tr(T^a G) for a Hermitian G is real; out[a] = 2 Tr(T^a G) = 2 Re Tr(T^a G)
template <typename R> __device__ __forceinline__ void two_re_trace_gen(const mat3<R> &w, R *out)
{
    const R s3 = R(0.57735026918962576450914878050196);
    out[0] = w.m[1].re + w.m[3].re;
    out[1] = w.m[1].im - w.m[3].im;
    out[2] = w.m[0].re - w.m[4].re;
    out[3] = w.m[2].re + w.m[6].re;
    out[4] = w.m[2].im - w.m[6].im;
    out[5] = w.m[5].re + w.m[7].re;
    out[6] = w.m[5].im - w.m[7].im;
    out[7] = s3 * (w.m[0].re + w.m[4].re - R(2) * w.m[8].re);
}

// Series length and scaling threshold of the exponential: Q is scaled so
// that its spectral radius is <= 1/4; the series then converges to below
// one unit in the last place after `terms` terms ((1/4)^13 / 13! = 2e-18
// for FP64, (1/4)^8 / 8! = 4e-10 for FP32).
template <typename R> struct exp_cfg;
template <> struct exp_cfg<double> {
    static constexpr int terms = 12;
};
template <> struct exp_cfg<float> {
    static constexpr int terms = 7;
};

// exp(iQ) for traceless Hermitian Q by Cayley-Hamilton:
//   Q^3 = c0 I + c1 Q,
// with c1 = tr(Q^2)/2, c0 = det Q, so exp(iQ) = f0 I + f1 Q + f2 Q^2 where the
// f's are the power series sum_n i^n (a_n, b_n, c_n)/n! of the recurrence
// (a,b,c)_{n+1} = (c0 c_n, a_n + c1 c_n, b_n).  Q is first scaled by 2^-s so
// that its spectral radius is <= 1/4 (|lambda|_max^2 <= 4 c1 / 3 for a
// traceless Hermitian 3x3), and the result is squared s times.
template <typename R> __device__ __forceinline__ mat3<R> exp_i_hermitian(mat3<R> q)
{
    // Q^2 is Hermitian: real diagonal, upper triangle, lower = conjugate
    mat3<R> q2;
    {
        const R d0 = q.m[0].re, d1 = q.m[4].re, d2 = q.m[8].re;
        const R a01 = q.m[1].re * q.m[1].re + q.m[1].im * q.m[1].im;
        const R a02 = q.m[2].re * q.m[2].re + q.m[2].im * q.m[2].im;
        const R a12 = q.m[5].re * q.m[5].re + q.m[5].im * q.m[5].im;
        q2.m[0] = {fmx(d0, d0, a01 + a02), R(0)};
        q2.m[4] = {fmx(d1, d1, a01 + a12), R(0)};
        q2.m[8] = {fmx(d2, d2, a02 + a12), R(0)};
        // (Q^2)_01 = q00 q01 + q01 q11 + q02 q21, q21 = conj(q12)
        cplx<R> t = {(d0 + d1) * q.m[1].re, (d0 + d1) * q.m[1].im};
        mac_nc(t, q.m[2], q.m[5]);
        q2.m[1] = t;
        // (Q^2)_02 = q00 q02 + q01 q12 + q02 q22
        t = {(d0 + d2) * q.m[2].re, (d0 + d2) * q.m[2].im};
        mac(t, q.m[1], q.m[5]);
        q2.m[2] = t;
        // (Q^2)_12 = q10 q02 + q11 q12 + q12 q22, q10 = conj(q01)
        t = {(d1 + d2) * q.m[5].re, (d1 + d2) * q.m[5].im};
        mac_cn(t, q.m[1], q.m[2]);
        q2.m[5] = t;
        q2.m[3] = conjg(q2.m[1]);
        q2.m[6] = conjg(q2.m[2]);
        q2.m[7] = conjg(q2.m[5]);
    }
    R c1 = R(0.5) * (q2.m[0].re + q2.m[4].re + q2.m[8].re);
    int s = 0;
    R rho2 = R(4) / R(3) * c1;
    R scale = R(1);
    while (rho2 > R(0.0625) && s < 64) { // bounded: a non-finite input cannot loop forever
        rho2 *= R(0.25);
        scale *= R(0.5);
        ++s;
    }
    if (s) {
#pragma unroll
        for (int k = 0; k < 9; ++k) q.m[k] = scale * q.m[k];
#pragma unroll
        for (int k = 0; k < 9; ++k) q2.m[k] = (scale * scale) * q2.m[k];
        c1 *= scale * scale;
    }
    // det Q, real for Hermitian Q:
    //   q00 q11 q22 + 2 Re(q01 q12 q20) - q00 |q12|^2 - q11 |q02|^2 - q22 |q01|^2
    R c0;
    {
        const R d0 = q.m[0].re, d1 = q.m[4].re, d2 = q.m[8].re;
        cplx<R> t = {R(0), R(0)};
        mac(t, q.m[1], q.m[5]); // q01 q12
        R tri = t.re * q.m[2].re + t.im * q.m[2].im; // Re(q01 q12 conj(q02))
        const R a01 = q.m[1].re * q.m[1].re + q.m[1].im * q.m[1].im;
        const R a02 = q.m[2].re * q.m[2].re + q.m[2].im * q.m[2].im;
        const R a12 = q.m[5].re * q.m[5].re + q.m[5].im * q.m[5].im;
        c0 = d0 * d1 * d2 + R(2) * tri - d0 * a12 - d1 * a02 - d2 * a01;
    }
    // series, fully unrolled with compile-time reciprocals
    R a = 1, b = 0, c = 0; // (a,b,c)_n / n!
    R f0r = 1, f0i = 0, f1r = 0, f1i = 0, f2r = 0, f2i = 0;
#pragma unroll
    for (int n = 0; n < exp_cfg<R>::terms; ++n) {
        const R inv = R(1) / R(n + 1);
        const R na = c0 * c, nb = fmx(c1, c, a), nc = b;
        a = na * inv;
        b = nb * inv;
        c = nc * inv;
        // multiply by i^(n+1)
        switch ((n + 1) & 3) {
        case 0: f0r += a; f1r += b; f2r += c; break;
        case 1: f0i += a; f1i += b; f2i += c; break;
        case 2: f0r -= a; f1r -= b; f2r -= c; break;
        default: f0i -= a; f1i -= b; f2i -= c; break;
        }
    }
    const cplx<R> f0 = {f0r, f0i}, f1 = {f1r, f1i}, f2 = {f2r, f2i};
    mat3<R> x;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        cplx<R> t = {R(0), R(0)};
        mac(t, f1, q.m[k]);
        mac(t, f2, q2.m[k]);
        x.m[k] = t;
    }
    x.m[0] = x.m[0] + f0;
    x.m[4] = x.m[4] + f0;
    x.m[8] = x.m[8] + f0;
    for (int k = 0; k < s; ++k) x = mul(x, x);
    return x;
}

// Host-visible lattice geometry, passed by value to kernels.
struct geom {
    int32_t n0, n1, n2;   // extents in x, y, z
    int64_t vol;          // n0 n1 n2
    double a[4];          // a0 (time), a1, a2, a3
};

// A lattice site with its coordinates: s = x + n0 (y + n1 z), 0-based, x
// fastest (lattice.f90:883-898 on 0-based indices).  Kernels derive the
// coordinates once per site (32-bit arithmetic; the library limits V to
// 2^31 - 1) and step to neighbours by comparison, never by division.
struct lsite {
    int64_t s;
    int x, y, z;
};

__device__ __forceinline__ lsite make_site(const geom &g, int64_t s)
{
    const unsigned su = (unsigned)s, n0 = (unsigned)g.n0, n1 = (unsigned)g.n1;
    const unsigned q = su / n0;
    const unsigned z = q / n1;
    return lsite{s, (int)(su - q * n0), (int)(q - z * n1), (int)z};
}

// Periodic neighbour p + d^ (d = 0, 1, 2), lattice.f90:846-874.
__device__ __forceinline__ lsite site_up(const geom &g, lsite p, int d)
{
    if (d == 0) {
        if (p.x == g.n0 - 1) {
            p.x = 0;
            p.s -= g.n0 - 1;
        } else {
            ++p.x;
            ++p.s;
        }
    } else if (d == 1) {
        if (p.y == g.n1 - 1) {
            p.y = 0;
            p.s -= (int64_t)(g.n1 - 1) * g.n0;
        } else {
            ++p.y;
            p.s += g.n0;
        }
    } else {
        const int64_t pl = (int64_t)g.n0 * g.n1;
        if (p.z == g.n2 - 1) {
            p.z = 0;
            p.s -= (int64_t)(g.n2 - 1) * pl;
        } else {
            ++p.z;
            p.s += pl;
        }
    }
    return p;
}

// Periodic neighbour p - d^.
__device__ __forceinline__ lsite site_dn(const geom &g, lsite p, int d)
{
    if (d == 0) {
        if (p.x == 0) {
            p.x = g.n0 - 1;
            p.s += g.n0 - 1;
        } else {
            --p.x;
            --p.s;
        }
    } else if (d == 1) {
        if (p.y == 0) {
            p.y = g.n1 - 1;
            p.s += (int64_t)(g.n1 - 1) * g.n0;
        } else {
            --p.y;
            p.s -= g.n0;
        }
    } else {
        const int64_t pl = (int64_t)g.n0 * g.n1;
        if (p.z == 0) {
            p.z = g.n2 - 1;
            p.s += (int64_t)(g.n2 - 1) * pl;
        } else {
            --p.z;
            p.s -= pl;
        }
    }
    return p;
}

} // namespace nrq
