// nrq_cnn.hip -- learned preconditioner for the CN systems: a 4-layer 3-D
// convolutional network N (18 -> 32 -> 32 -> 32 -> 18 real channels, 3^3
// periodic kernels, ReLU between layers, no biases) acting on one colour
// vector field v and its covariant features f1 = beta K v, f2 = beta K f1
// (3 features x 3 colours x re/im = 18 channels per site) of the CN operator
// A = 1 + beta K.  The preconditioner is z = v - f1 + f2 + N(v, f1, f2): the
// second-order Neumann polynomial of A^-1 corrected by the network (all
// weights 0: the polynomial alone), and training fits N(r, ...) to
// e - (r - f1 + f2) on residual/error pairs (r, e), r = A e, harvested from
// plain GMRES solves.  N is positively homogeneous (ReLU, no biases), so it
// commutes with the scale of its input.
//
// Convolutions are implicit GEMMs on BF16 matrix cores with FP32
// accumulation: a wavefront computes eight runs of 16 consecutive sites x 32
// output channels with one v_mfma_f32_16x16x32_bf16 per (run, tap, 16 output
// channels), each tap's weights loaded once for the eight runs;
// K = 32 input channels per tap.  Activations are channels-last BF16
// [site][32] (channels past 18 are zero in the outer layers), weights
// [layer][tap][co][ci] with tap = 9 (dz+1) + 3 (dy+1) + (dx+1).  Backward:
// d(input) is the same convolution with mirrored, transposed weights; the
// weight gradient is a GEMM over sites (d(pre-activation) x shifted input) on
// the same matrix cores.  Adam updates FP32 master weights.
#include "nrq_cnn.h"
#include "nrq_krylov.h"
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

using namespace nrq;

namespace {

constexpr int CH = 32, TAPS = 27, NL = 4, CIN = 18;
constexpr int WL = TAPS * CH * CH;   // weights per layer (padded)
constexpr int NW = NL * WL;          // all weights
constexpr int DW_SITES = 4096;       // sites per block of the weight-gradient kernel

typedef uint16_t bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline bf16 f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u); // round to nearest even
    return (bf16)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16 h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__device__ __forceinline__ int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// out[s][co] = sum_{tap, ci} w[tap][co][ci] in[s + off(tap)][ci] for one
// layer.  out_f != null: FP32 output; else BF16 output, with ReLU when relu,
// and zeroed where mask[s][co] <= 0 when mask != null (the ReLU derivative of
// the backward pass).  Grid: 4 wavefronts per block; a wavefront computes
// CONV_SEG consecutive 16-site x runs (in x, then y, then z order) and loads
// each tap's weights once for all of them.
constexpr int CONV_SEG = 8;
__global__ __launch_bounds__(256) void k_conv(const bf16 *__restrict__ in, const bf16 *__restrict__ w,
                                              bf16 *__restrict__ out_b, float *__restrict__ out_f,
                                              const bf16 *__restrict__ mask, int relu, geom g)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nsx = g.n0 >> 4;
    const int64_t nseg = (int64_t)nsx * g.n1 * g.n2;
    const int64_t seg0 = ((int64_t)blockIdx.x * 4 + wv) * CONV_SEG;
    if (seg0 >= nseg) return;
    const int m = lane & 15, kq = lane >> 4;
    int xs[CONV_SEG], ys[CONV_SEG], zs[CONV_SEG];
#pragma unroll
    for (int j = 0; j < CONV_SEG; ++j) {
        const int64_t sg = seg0 + j < nseg ? seg0 + j : seg0; // a run past the end repeats the first
        xs[j] = (int)(sg % nsx);
        const int64_t yz = sg / nsx;
        ys[j] = (int)(yz % g.n1);
        zs[j] = (int)(yz / g.n1);
    }
    f32x4 acc0[CONV_SEG], acc1[CONV_SEG];
#pragma unroll
    for (int j = 0; j < CONV_SEG; ++j) acc0[j] = acc1[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    int tap = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx, ++tap) {
                // B[k][n] = w[tap][co n][ci k], shared by the runs
                const bf16x8 b0 = *(const bf16x8 *)(w + ((int64_t)tap * CH + m) * CH + 8 * kq);
                const bf16x8 b1 = *(const bf16x8 *)(w + ((int64_t)tap * CH + 16 + m) * CH + 8 * kq);
#pragma unroll
                for (int j = 0; j < CONV_SEG; ++j) {
                    const int xx = wrap(xs[j] * 16 + m + dx, g.n0);
                    const int yy = wrap(ys[j] + dy, g.n1), zz = wrap(zs[j] + dz, g.n2);
                    const int64_t nb = xx + (int64_t)g.n0 * (yy + (int64_t)g.n1 * zz);
                    // A[m][k] = in[site m][ci k]
                    const bf16x8 a = *(const bf16x8 *)(in + nb * CH + 8 * kq);
                    acc0[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc0[j], 0, 0, 0);
                    acc1[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc1[j], 0, 0, 0);
                }
            }
    // lane holds D[site 4 kq + i][co m] and D[site 4 kq + i][co 16 + m] of each run
#pragma unroll
    for (int j = 0; j < CONV_SEG; ++j) {
        if (seg0 + j >= nseg) break;
        const int64_t base = (int64_t)xs[j] * 16 + (int64_t)g.n0 * (ys[j] + (int64_t)g.n1 * zs[j]);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t s = base + 4 * kq + i;
            float v0 = acc0[j][i], v1 = acc1[j][i];
            if (out_f) {
                out_f[s * CH + m] = v0;
                out_f[s * CH + 16 + m] = v1;
            } else {
                if (relu) {
                    v0 = fmaxf(v0, 0.f);
                    v1 = fmaxf(v1, 0.f);
                }
                if (mask) {
                    if (!(bf2f(mask[s * CH + m]) > 0.f)) v0 = 0.f;
                    if (!(bf2f(mask[s * CH + 16 + m]) > 0.f)) v1 = 0.f;
                }
                out_b[s * CH + m] = f2bf(v0);
                out_b[s * CH + 16 + m] = f2bf(v1);
            }
        }
    }
}

// k_conv with the input staged in LDS, for lattices with n0 % 128 == 0, even
// n1 and n2 % ZC == 0: a block of 4 wavefronts owns a 128 x 2 tile of (x, y)
// and sweeps ZC planes in z.  LDS holds the layer's weights in the lane order
// of the MFMA B operands (55 KB, read once per block instead of from L2 at
// every tap) and a ring of three input planes of the tile with a one-site
// halo (130 x 4 sites x 32 BF16 channels each, 100 KB);
// output plane z reads planes z - 1, z, z + 1 there instead of re-reading each
// input site 27 times through L1, while plane z + 2 streams into registers
// and then replaces plane z - 1.  Wavefront w computes row y0 + (w & 1), x runs
// 4 (w >> 1) .. 4 (w >> 1) + 3.  Same arguments, same arithmetic and summation
// order per output as k_conv.
constexpr int LT_X = 128, LT_HX = LT_X + 2, LT_HY = 4;
constexpr int LT_PLANE = LT_HX * LT_HY * 4;                     // bf16x8 per plane
constexpr int LT_PER = (LT_PLANE + 255) / 256;                  // per thread and plane
template <int ZC>
__global__ __launch_bounds__(256) void k_conv_lds(const bf16 *__restrict__ in, const bf16 *__restrict__ w,
                                                  bf16 *__restrict__ out_b, float *__restrict__ out_f,
                                                  const bf16 *__restrict__ mask, int relu, geom g)
{
    __shared__ bf16x8 lds[3 * LT_PLANE];
    // the layer's weights as the MFMA B operands: [tap][half][lane]
    __shared__ bf16x8 wl[27 * 2 * 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int q = threadIdx.x; q < 27 * 2 * 64; q += 256) {
        const int l = q & 63, hf = (q >> 6) & 1, tp = q >> 7;
        wl[q] = *(const bf16x8 *)(w + ((int64_t)tp * CH + 16 * hf + (l & 15)) * CH + 8 * (l >> 4));
    }
    const int ntx = g.n0 / LT_X, nty = g.n1 / 2;
    const int b = blockIdx.x;
    const int x0 = (b % ntx) * LT_X, y0 = ((b / ntx) % nty) * 2, z0 = (b / (ntx * nty)) * ZC;
    bf16x8 pre[LT_PER];
    auto gload = [&](int z) {
        const int gz = wrap(z, g.n2);
#pragma unroll
        for (int k = 0; k < LT_PER; ++k) {
            const int q = threadIdx.x + 256 * k;
            if (q < LT_PLANE) {
                const int c4 = q & 3, site = q >> 2;
                const int hx = site % LT_HX, hy = site / LT_HX;
                const int gx = wrap(x0 - 1 + hx, g.n0), gy = wrap(y0 - 1 + hy, g.n1);
                const int64_t gs = gx + (int64_t)g.n0 * (gy + (int64_t)g.n1 * gz);
                pre[k] = *(const bf16x8 *)(in + gs * CH + 8 * c4);
            }
        }
    };
    auto sstore = [&](int slot) {
#pragma unroll
        for (int k = 0; k < LT_PER; ++k) {
            const int q = threadIdx.x + 256 * k;
            if (q < LT_PLANE) lds[slot * LT_PLANE + q] = pre[k];
        }
    };
    for (int p = 0; p < 3; ++p) {
        gload(z0 - 1 + p);
        sstore(p);
    }
    __syncthreads();
    const int m = lane & 15, kq = lane >> 4;
    const int ry = wv & 1, xr0 = 4 * (wv >> 1);
#pragma unroll 1
    for (int kz = 0; kz < ZC; ++kz) {
        const bool more = kz + 1 < ZC;
        if (more) gload(z0 + kz + 2);
        f32x4 acc0[4], acc1[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) acc0[j] = acc1[j] = f32x4{0.f, 0.f, 0.f, 0.f};
        int tap = 0;
        for (int dz = -1; dz <= 1; ++dz) {
            const int slot = (kz + 1 + dz) % 3; // slots hold planes z0 - 1 + p at p mod 3
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx, ++tap) {
                    const bf16x8 b0 = wl[(tap * 2 + 0) * 64 + lane];
                    const bf16x8 b1 = wl[(tap * 2 + 1) * 64 + lane];
                    const bf16x8 *row = lds + slot * LT_PLANE + ((ry + dy + 1) * LT_HX + 16 * xr0 + m + dx + 1) * 4 + kq;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const bf16x8 a = row[16 * j * 4];
                        acc0[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc0[j], 0, 0, 0);
                        acc1[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc1[j], 0, 0, 0);
                    }
                }
        }
        const int64_t base = x0 + 16 * xr0 + (int64_t)g.n0 * ((y0 + ry) + (int64_t)g.n1 * (z0 + kz));
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t s = base + 16 * j + 4 * kq + i;
                float v0 = acc0[j][i], v1 = acc1[j][i];
                if (out_f) {
                    out_f[s * CH + m] = v0;
                    out_f[s * CH + 16 + m] = v1;
                } else {
                    if (relu) {
                        v0 = fmaxf(v0, 0.f);
                        v1 = fmaxf(v1, 0.f);
                    }
                    if (mask) {
                        if (!(bf2f(mask[s * CH + m]) > 0.f)) v0 = 0.f;
                        if (!(bf2f(mask[s * CH + 16 + m]) > 0.f)) v1 = 0.f;
                    }
                    out_b[s * CH + m] = f2bf(v0);
                    out_b[s * CH + 16 + m] = f2bf(v1);
                }
            }
        __syncthreads(); // plane z - 1 is no longer read
        if (more) sstore(kz % 3);
        __syncthreads();
    }
}

// Partial weight gradient of one block of DW_SITES sites, on matrix cores:
//   part[blk][tap][co][ci] = sum_s d[s][co] in[s + off(tap)][ci]
// a GEMM over sites, 32 per MFMA step: A[co][site] = d, B[site][ci] = the
// shifted input.  Wavefront wv accumulates the taps wv, wv + 4, ... (<= 7).
__global__ __launch_bounds__(256) void k_dw(const bf16 *__restrict__ d, const bf16 *__restrict__ in,
                                            float *__restrict__ part, geom g)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int m = lane & 15, kq = lane >> 4;
    f32x4 acc[7][2][2];
#pragma unroll
    for (int j = 0; j < 7; ++j)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[j][a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int64_t s0 = (int64_t)blockIdx.x * DW_SITES;
    for (int c0 = 0; c0 < DW_SITES; c0 += 32) {
        // this lane's 8 sites of the 32-site step and their coordinates
        int64_t sv[8];
        int xs[8], ys[8], zs[8];
        bool ok[8];
        bf16x8 a0, a1;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int64_t s = s0 + c0 + 8 * kq + i;
            ok[i] = s < g.vol;
            sv[i] = ok[i] ? s : 0;
            xs[i] = (int)(sv[i] % g.n0);
            const int64_t yz = sv[i] / g.n0;
            ys[i] = (int)(yz % g.n1);
            zs[i] = (int)(yz / g.n1);
            const uint16_t h0 = ok[i] ? d[sv[i] * CH + m] : (uint16_t)0, h1 = ok[i] ? d[sv[i] * CH + 16 + m] : (uint16_t)0;
            a0[i] = __builtin_bit_cast(__bf16, h0);
            a1[i] = __builtin_bit_cast(__bf16, h1);
        }
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            const int tap = wv + 4 * j;
            if (tap >= TAPS) break;
            const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
            bf16x8 b0, b1;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int64_t nb = wrap(xs[i] + dx, g.n0) + (int64_t)g.n0 * (wrap(ys[i] + dy, g.n1) + (int64_t)g.n1 * wrap(zs[i] + dz, g.n2));
                const uint16_t h0 = ok[i] ? in[nb * CH + m] : (uint16_t)0, h1 = ok[i] ? in[nb * CH + 16 + m] : (uint16_t)0;
                b0[i] = __builtin_bit_cast(__bf16, h0);
                b1[i] = __builtin_bit_cast(__bf16, h1);
            }
            acc[j][0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[j][0][0], 0, 0, 0);
            acc[j][0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[j][0][1], 0, 0, 0);
            acc[j][1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[j][1][0], 0, 0, 0);
            acc[j][1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[j][1][1], 0, 0, 0);
        }
    }
    // lane holds C[co = 16 a + 4 kq + i][ci = 16 b + m]
    float *p = part + (int64_t)blockIdx.x * WL;
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        const int tap = wv + 4 * j;
        if (tap >= TAPS) break;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    p[((int64_t)tap * CH + 16 * a + 4 * kq + i) * CH + 16 * b + m] = acc[j][a][b][i];
    }
}

// out[i] = sum_b part[b][i] (fixed order), i < WL
__global__ void k_dw_fold(const float *__restrict__ part, int nblk, float *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= WL) return;
    float s = 0.f;
    for (int b = 0; b < nblk; ++b) s += part[(int64_t)b * WL + i];
    out[i] = s;
}

// Loss sum_{s, c < 18} (o[s][c] - t[s][c])^2 / norm: per-block partials and
// the gradient d[s][c] = 2 (o - t) / norm in BF16 (channels >= 18: 0).
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ o, const float *__restrict__ t, double inv_norm,
                                              bf16 *__restrict__ d, double *__restrict__ part, int64_t vol)
{
    __shared__ double sh[4];
    double acc = 0;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < vol; s += (int64_t)gridDim.x * blockDim.x) {
        for (int c = 0; c < CH; ++c) {
            float g = 0.f;
            if (c < CIN) {
                const double e = (double)o[s * CH + c] - (double)t[s * CIN + c];
                acc += e * e;
                g = (float)(2.0 * e * inv_norm);
            }
            d[s * CH + c] = f2bf(g);
        }
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1] + sh[2] + sh[3]) * inv_norm;
}

__global__ void k_sum(const double *__restrict__ part, int n, double *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0;
    for (int i = 0; i < n; ++i) s += part[i];
    *out += s;
}

// Adam on the FP32 master weights; also writes the BF16 weights and the
// mirrored-transposed BF16 copy used by the backward convolution.
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1, float *__restrict__ m2,
                       float lr, float b1, float b2, float c1, float c2, bf16 *__restrict__ wb, bf16 *__restrict__ wt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NW) return;
    if (g) {
        const float gi = g[i];
        m1[i] = b1 * m1[i] + (1.f - b1) * gi;
        m2[i] = b2 * m2[i] + (1.f - b2) * gi * gi;
        w[i] -= lr * (m1[i] * c1) / (sqrtf(m2[i] * c2) + 1e-8f);
    }
    const int l = i / WL, r = i % WL, tap = r / (CH * CH), co = (r / CH) % CH, ci = r % CH;
    wb[i] = f2bf(w[i]);
    wt[(int64_t)l * WL + ((int64_t)(26 - tap) * CH + ci) * CH + co] = f2bf(w[i]);
}

// Complex scale factors of one column with CN coefficient beta and kmax =
// sum_i 2 / a_i^2 (the top of the spectrum of K): the network's inputs are v,
// f1 s1 = K v / kmax and f2 s2 = K^2 v / kmax^2, and its output is the
// remainder of the Neumann polynomial divided by c = -(beta kmax)^3, whose
// leading term is c (K / kmax)^3 v.  Inputs and target are then the same
// functions of v whatever the mass (beta).
struct col_scale {
    cplx<double> s1, s2, c, cinv;
};
__host__ inline col_scale make_col_scale(cplx<double> beta, double kmax)
{
    const std::complex<double> b(beta.re * kmax, beta.im * kmax);
    const std::complex<double> i1 = 1.0 / b, i2 = i1 * i1, c = -(b * b * b), ci = 1.0 / c;
    return {{i1.real(), i1.imag()}, {i2.real(), i2.imag()}, {c.real(), c.imag()}, {ci.real(), ci.imag()}};
}
__device__ __forceinline__ cplx<double> cmuld(cplx<double> a, cplx<double> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// BF16 network input of one colour field v (complex colour-major, 3V) and
// its features f1 = beta K v, f2 = beta K f1: channel 6 q + 2 c + {re, im}
// of feature q (v, f1 s1, f2 s2) and colour c; channels >= 18 zero.
template <typename T>
__global__ void k_pack(const cplx<T> *__restrict__ v, const cplx<T> *__restrict__ f1, const cplx<T> *__restrict__ f2,
                       bf16 *__restrict__ act, int64_t vol, col_scale sc)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < vol; s += (int64_t)gridDim.x * blockDim.x) {
        const cplx<T> *f[3] = {v, f1, f2};
        const cplx<double> m[3] = {{1.0, 0.0}, sc.s1, sc.s2};
        for (int q = 0; q < 3; ++q)
            for (int c = 0; c < 3; ++c) {
                const cplx<T> a0 = f[q][(int64_t)c * vol + s];
                const cplx<double> a = cmuld({(double)a0.re, (double)a0.im}, m[q]);
                act[s * CH + 6 * q + 2 * c] = f2bf((float)a.re);
                act[s * CH + 6 * q + 2 * c + 1] = f2bf((float)a.im);
            }
        for (int c = CIN; c < CH; ++c) act[s * CH + c] = 0;
    }
}

// z = v - f1 + f2 + c N(v, f1 s1, f2 s2): the second-order Neumann
// polynomial of (1 + beta K)^-1 plus the network's correction, output
// channels 0..5 (colour c at 2 c, 2 c + 1) scaled by c = -(beta kmax)^3
template <typename T>
__global__ void k_unpack(const float *__restrict__ o, const cplx<T> *__restrict__ v, const cplx<T> *__restrict__ f1,
                         const cplx<T> *__restrict__ f2, cplx<T> *__restrict__ z, int64_t vol, col_scale sc)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < vol; s += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < 3; ++c) {
            const int64_t at = (int64_t)c * vol + s;
            const cplx<T> a = v[at], b = f1[at], d = f2[at];
            const cplx<double> n = cmuld({(double)o[s * CH + 2 * c], (double)o[s * CH + 2 * c + 1]}, sc.c);
            z[at] = {(T)((a.re - b.re + d.re) + n.re), (T)((a.im - b.im + d.im) + n.im)};
        }
}

// Training sample of one column of a CN solve: right-hand side y, solution
// x = A^-1 y, the features f1 = beta K y, f2 = beta K f1 and the column's
// scales (col_scale); input [y, f1 s1, f2 s2] / |y| (BF16), target
// (x - (y - f1 + f2)) / (c |y|) (what the second-order Neumann polynomial
// misses, in units of its leading term's scale) in channels 0..5 and zero in
// 6..17 (FP32 [V][18]).
__global__ void k_make_sample(const cplx<double> *__restrict__ y, const cplx<double> *__restrict__ x,
                              const cplx<double> *__restrict__ f1, const cplx<double> *__restrict__ f2, double inv,
                              col_scale sc, bf16 *__restrict__ act, float *__restrict__ tgt, int64_t vol)
{
    const cplx<double> cinv = {sc.cinv.re * inv, sc.cinv.im * inv};
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < vol; s += (int64_t)gridDim.x * blockDim.x) {
        const cplx<double> *f[3] = {y, f1, f2};
        const cplx<double> m[3] = {{inv, 0.0}, {sc.s1.re * inv, sc.s1.im * inv}, {sc.s2.re * inv, sc.s2.im * inv}};
        for (int q = 0; q < 3; ++q)
            for (int c = 0; c < 3; ++c) {
                const cplx<double> a = cmuld(f[q][(int64_t)c * vol + s], m[q]);
                act[s * CH + 6 * q + 2 * c] = f2bf((float)a.re);
                act[s * CH + 6 * q + 2 * c + 1] = f2bf((float)a.im);
            }
        for (int c = CIN; c < CH; ++c) act[s * CH + c] = 0;
        for (int c = 0; c < 3; ++c) {
            const int64_t at = (int64_t)c * vol + s;
            const cplx<double> r = {x[at].re - (y[at].re - f1[at].re + f2[at].re),
                                    x[at].im - (y[at].im - f1[at].im + f2[at].im)};
            const cplx<double> t = cmuld(r, cinv);
            tgt[s * CIN + 2 * c] = (float)t.re;
            tgt[s * CIN + 2 * c + 1] = (float)t.im;
        }
        for (int c = 6; c < CIN; ++c) tgt[s * CIN + c] = 0.f;
    }
}

// Gaussian random fields from a counter-based hash (independent of ranlxd)
__device__ inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__global__ void k_gauss(cplx<double> *__restrict__ f, int64_t n, uint64_t seed)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t a = mix64(seed * 0x9e3779b97f4a7c15ull + 2 * (uint64_t)i);
        const uint64_t b = mix64(seed * 0x9e3779b97f4a7c15ull + 2 * (uint64_t)i + 1);
        const double u1 = ((a >> 11) + 1) * (1.0 / 9007199254740993.0), u2 = (b >> 11) * (1.0 / 9007199254740992.0);
        const double r = sqrt(-2.0 * log(u1));
        f[i] = {r * cos(6.283185307179586 * u2), r * sin(6.283185307179586 * u2)};
    }
}

} // namespace

struct cnn_sample {
    bf16 *act = nullptr; // [V][32] input
    float *tgt = nullptr; // [V][18] target correction
};

struct nrq_cnn {
    std::vector<float> w_host;            // master weights (host copy)
    float *w = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
    bf16 *wb = nullptr, *wt = nullptr;    // BF16 weights and mirrored-transposed copy
    int64_t adam_t = 0;
    // per-geometry buffers
    geom g{};
    bf16 *act[NL] = {};                   // layer inputs (act[0] = network input)
    bf16 *d0 = nullptr, *d1 = nullptr;
    float *out = nullptr;                 // [V][32] FP32 network output
    float *part = nullptr;                // weight-gradient partials
    double *lpart = nullptr, *loss = nullptr;
    // training set (on its harvest geometry)
    geom data_g{};
    std::vector<cnn_sample> data;
};

namespace nrq {

static nrq_cnn *g_precond = nullptr;
nrq_cnn *cn_preconditioner() { return g_precond; }

static void free_geom_buffers(nrq_cnn *n)
{
    for (auto &a : n->act)
        if (a) (void)hipFree(a), a = nullptr;
    void *p[] = {n->d0, n->d1, n->out, n->part, n->lpart, n->loss};
    for (void *q : p)
        if (q) (void)hipFree(q);
    n->d0 = n->d1 = nullptr;
    n->out = nullptr;
    n->part = nullptr;
    n->lpart = n->loss = nullptr;
    n->g = geom{};
}

static int ensure_geom(nrq_cnn *n, const geom &g)
{
    if (n->g.vol == g.vol && n->g.n0 == g.n0 && n->g.n1 == g.n1 && n->g.n2 == g.n2) return NRQ_OK;
    if (g.n0 % 16) return fail(NRQ_ERR_UNSUPPORTED, "cnn: the x extent %d must be a multiple of 16", g.n0);
    free_geom_buffers(n);
    const size_t ab = (size_t)g.vol * CH * sizeof(bf16);
    const int nblk = (int)((g.vol + DW_SITES - 1) / DW_SITES);
    bool ok = true;
    for (auto &a : n->act) ok = ok && hipMalloc(&a, ab) == hipSuccess;
    ok = ok && hipMalloc(&n->d0, ab) == hipSuccess && hipMalloc(&n->d1, ab) == hipSuccess &&
         hipMalloc(&n->out, (size_t)g.vol * CH * sizeof(float)) == hipSuccess &&
         hipMalloc(&n->part, (size_t)nblk * WL * sizeof(float)) == hipSuccess &&
         hipMalloc(&n->lpart, 1024 * sizeof(double)) == hipSuccess && hipMalloc(&n->loss, sizeof(double)) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        free_geom_buffers(n);
        return fail(NRQ_ERR_ALLOC, "cnn: buffer allocation failed (V=%lld)", (long long)g.vol);
    }
    n->g = g;
    return NRQ_OK;
}

static unsigned conv_grid(const geom &g)
{
    const int64_t waves = ((int64_t)(g.n0 / 16) * g.n1 * g.n2 + CONV_SEG - 1) / CONV_SEG;
    return (unsigned)((waves + 3) / 4);
}

// One convolution layer (k_conv arguments), LDS-staged where the extents
// allow it.
static void conv(const geom &g, const bf16 *in, const bf16 *w, bf16 *out_b, float *out_f, const bf16 *mask, int relu)
{
    if (g.n0 % LT_X == 0 && g.n1 % 2 == 0 && g.n2 % 16 == 0) {
        const unsigned blocks = (unsigned)((g.n0 / LT_X) * (g.n1 / 2) * (g.n2 / 16));
        hipLaunchKernelGGL(k_conv_lds<16>, dim3(blocks), dim3(256), 0, stream(), in, w, out_b, out_f, mask, relu, g);
    } else if (g.n0 % LT_X == 0 && g.n1 % 2 == 0 && g.n2 % 8 == 0) {
        const unsigned blocks = (unsigned)((g.n0 / LT_X) * (g.n1 / 2) * (g.n2 / 8));
        hipLaunchKernelGGL(k_conv_lds<8>, dim3(blocks), dim3(256), 0, stream(), in, w, out_b, out_f, mask, relu, g);
    } else {
        hipLaunchKernelGGL(k_conv, dim3(conv_grid(g)), dim3(256), 0, stream(), in, w, out_b, out_f, mask, relu, g);
    }
}

// forward pass from n->act[0]; FP32 output in n->out
static int forward(nrq_cnn *n)
{
    const geom &g = n->g;
    for (int l = 0; l < NL; ++l) {
        const bf16 *w = n->wb + (int64_t)l * WL;
        if (l < NL - 1)
            conv(g, n->act[l], w, n->act[l + 1], nullptr, nullptr, 1);
        else
            conv(g, n->act[l], w, nullptr, n->out, nullptr, 0);
    }
    NRQ_HIP(hipGetLastError());
    return NRQ_OK;
}

// loss of the current output against tgt (added to n->loss) and the weight
// gradient into n->grad; with dtrace (host, NL x V x 32 bf16), the
// gradient at the output of each layer l (the d of its weight gradient) is
// copied to dtrace[l]
static int backward(nrq_cnn *n, const float *tgt, double norm, bf16 *dtrace = nullptr)
{
    const geom &g = n->g;
    const int nblk = (int)((g.vol + DW_SITES - 1) / DW_SITES);
    hipLaunchKernelGGL(k_loss, dim3(512), dim3(256), 0, stream(), (const float *)n->out, tgt, 1.0 / norm, n->d0, n->lpart, g.vol);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(64), 0, stream(), (const double *)n->lpart, 512, n->loss);
    bf16 *d = n->d0, *dn = n->d1;
    for (int l = NL - 1; l >= 0; --l) {
        if (dtrace)
            NRQ_HIP(hipMemcpyAsync(dtrace + (size_t)l * g.vol * CH, d, sizeof(bf16) * (size_t)g.vol * CH,
                                   hipMemcpyDeviceToHost, stream()));
        hipLaunchKernelGGL(k_dw, dim3(nblk), dim3(256), 0, stream(), (const bf16 *)d, (const bf16 *)n->act[l], n->part, g);
        hipLaunchKernelGGL(k_dw_fold, dim3((WL + 255) / 256), dim3(256), 0, stream(), (const float *)n->part, nblk,
                           n->grad + (int64_t)l * WL);
        if (l > 0) {
            // d(pre-activation of layer l-1) = conv(d, mirrored W_l^T) where act[l] > 0
            conv(g, d, n->wt + (int64_t)l * WL, dn, nullptr, n->act[l], 0);
            std::swap(d, dn);
        }
    }
    NRQ_HIP(hipGetLastError());
    return NRQ_OK;
}

static int upload_weights(nrq_cnn *n)
{
    NRQ_HIP(hipMemcpyAsync(n->w, n->w_host.data(), sizeof(float) * NW, hipMemcpyHostToDevice, stream()));
    hipLaunchKernelGGL(k_adam, dim3((NW + 255) / 256), dim3(256), 0, stream(), n->w, (const float *)nullptr, n->m1, n->m2, 0.f,
                       0.f, 0.f, 0.f, 0.f, n->wb, n->wt);
    NRQ_HIP(hipGetLastError());
    return NRQ_OK;
}

static double spectrum_top(const geom &g)
{
    return 2.0 / (g.a[1] * g.a[1]) + 2.0 / (g.a[2] * g.a[2]) + 2.0 / (g.a[3] * g.a[3]);
}

template <typename T>
int cnn_precondition(nrq_cnn *net, const cplx<T> *const *v, const cplx<T> *const *f1, const cplx<T> *const *f2,
                     cplx<T> *const *z, int R, const int *act, const cplx<double> *beta)
{
    const geom g = lat().g;
    if (int e = ensure_geom(net, g)) return e;
    const double kmax = spectrum_top(g);
    for (int r = 0; r < R; ++r) {
        if (act && !act[r]) continue;
        const col_scale sc = make_col_scale(beta[r], kmax);
        hipLaunchKernelGGL(k_pack<T>, dim3(1024), dim3(256), 0, stream(), v[r], f1[r], f2[r], net->act[0], g.vol, sc);
        if (int e = forward(net)) return e;
        hipLaunchKernelGGL(k_unpack<T>, dim3(1024), dim3(256), 0, stream(), (const float *)net->out, v[r], f1[r], f2[r],
                           z[r], g.vol, sc);
        NRQ_HIP(hipGetLastError());
    }
    return NRQ_OK;
}
template int cnn_precondition<double>(nrq_cnn *, const cplx<double> *const *, const cplx<double> *const *,
                                      const cplx<double> *const *, cplx<double> *const *, int, const int *,
                                      const cplx<double> *);
template int cnn_precondition<float>(nrq_cnn *, const cplx<float> *const *, const cplx<float> *const *,
                                     const cplx<float> *const *, cplx<float> *const *, int, const int *,
                                     const cplx<double> *);

} // namespace nrq

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

#define NEED_NET(n)                                                       \
    do {                                                                  \
        if (!(n)) return fail(NRQ_ERR_ARG, "cnn: null network");         \
    } while (0)

extern "C" int nrq_cnn_create(nrq_cnn **out, uint64_t seed)
{
    if (!out) return fail(NRQ_ERR_ARG, "cnn: null output handle");
    if (int r = ensure_device()) return r;
    nrq_cnn *n = new nrq_cnn();
    n->w_host.assign(NW, 0.f);
    // He initialisation of the real channels; the correction starts small
    // (scale 0.1 on the last layer) so that the untrained preconditioner is
    // close to the identity
    uint64_t st = seed * 0x9e3779b97f4a7c15ull + 12345;
    auto uni = [&]() {
        st += 0x9e3779b97f4a7c15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < NL; ++l) {
        const int cin = l == 0 ? CIN : CH, cout = l == NL - 1 ? CIN : CH;
        const double sd = std::sqrt(2.0 / (27.0 * cin)) * (l == NL - 1 ? 0.1 : 1.0);
        for (int tap = 0; tap < TAPS; ++tap)
            for (int co = 0; co < cout; ++co)
                for (int ci = 0; ci < cin; ++ci) {
                    const double g = std::sqrt(-2.0 * std::log(uni())) * std::cos(6.283185307179586 * uni());
                    n->w_host[(size_t)l * WL + ((size_t)tap * CH + co) * CH + ci] = (float)(sd * g);
                }
    }
    bool ok = hipMalloc(&n->w, sizeof(float) * NW) == hipSuccess && hipMalloc(&n->m1, sizeof(float) * NW) == hipSuccess &&
              hipMalloc(&n->m2, sizeof(float) * NW) == hipSuccess && hipMalloc(&n->grad, sizeof(float) * NW) == hipSuccess &&
              hipMalloc(&n->wb, sizeof(bf16) * NW) == hipSuccess && hipMalloc(&n->wt, sizeof(bf16) * NW) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        delete n;
        return fail(NRQ_ERR_ALLOC, "cnn: allocation failed");
    }
    NRQ_HIP(hipMemsetAsync(n->m1, 0, sizeof(float) * NW, stream()));
    NRQ_HIP(hipMemsetAsync(n->m2, 0, sizeof(float) * NW, stream()));
    if (int r = upload_weights(n)) return r;
    NRQ_HIP(hipStreamSynchronize(stream()));
    *out = n;
    return NRQ_OK;
}

extern "C" int nrq_cnn_destroy(nrq_cnn *n)
{
    if (!n) return NRQ_OK;
    if (stream()) (void)hipStreamSynchronize(stream());
    if (g_precond == n) g_precond = nullptr;
    free_geom_buffers(n);
    void *p[] = {n->w, n->m1, n->m2, n->grad, n->wb, n->wt};
    for (void *q : p)
        if (q) (void)hipFree(q);
    for (auto &s : n->data) {
        if (s.act) (void)hipFree(s.act);
        if (s.tgt) (void)hipFree(s.tgt);
    }
    delete n;
    return NRQ_OK;
}

extern "C" int nrq_cnn_get_weights(const nrq_cnn *n, float *w)
{
    NEED_NET(n);
    if (!w) return fail(NRQ_ERR_ARG, "cnn: null argument");
    NRQ_HIP(hipMemcpyAsync(w, n->w, sizeof(float) * NW, hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    return NRQ_OK;
}

extern "C" int nrq_cnn_set_weights(nrq_cnn *n, const float *w)
{
    NEED_NET(n);
    if (!w) return fail(NRQ_ERR_ARG, "cnn: null argument");
    memcpy(n->w_host.data(), w, sizeof(float) * NW);
    if (int r = upload_weights(n)) return r;
    NRQ_HIP(hipStreamSynchronize(stream()));
    return NRQ_OK;
}

extern "C" int nrq_cnn_forward(nrq_cnn *n, const float *in, float *out)
{
    NEED_NET(n);
    if (!lat().initialised) return fail(NRQ_ERR_STATE, "lattice: not initialised");
    if (!in || !out) return fail(NRQ_ERR_ARG, "cnn: null argument");
    const geom g = lat().g;
    if (int e = ensure_geom(n, g)) return e;
    std::vector<bf16> a((size_t)g.vol * CH, 0);
    for (int64_t s = 0; s < g.vol; ++s)
        for (int c = 0; c < CIN; ++c) a[(size_t)s * CH + c] = f2bf(in[(size_t)s * CIN + c]);
    NRQ_HIP(hipMemcpyAsync(n->act[0], a.data(), a.size() * sizeof(bf16), hipMemcpyHostToDevice, stream()));
    if (int e = forward(n)) return e;
    std::vector<float> o((size_t)g.vol * CH);
    NRQ_HIP(hipMemcpyAsync(o.data(), n->out, o.size() * sizeof(float), hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    for (int64_t s = 0; s < g.vol; ++s)
        for (int c = 0; c < CIN; ++c) out[(size_t)s * CIN + c] = o[(size_t)s * CH + c];
    return NRQ_OK;
}

extern "C" int nrq_cnn_time_forward(nrq_cnn *n, int reps, double *ms)
{
    NEED_NET(n);
    if (!lat().initialised) return fail(NRQ_ERR_STATE, "lattice: not initialised");
    if (!ms || reps < 1) return fail(NRQ_ERR_ARG, "cnn: bad argument");
    const geom g = lat().g;
    if (int e = ensure_geom(n, g)) return e;
    // the time of the matrix-core passes does not depend on the values
    NRQ_HIP(hipMemsetAsync(n->act[0], 0, (size_t)g.vol * CH * sizeof(bf16), stream()));
    hipEvent_t ev[2];
    NRQ_HIP(hipEventCreate(&ev[0]));
    if (hipEventCreate(&ev[1]) != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        return fail(NRQ_ERR_HIP, "cnn: event creation failed");
    }
    int rc = forward(n); // warm-up
    if (!rc && hipEventRecord(ev[0], stream()) != hipSuccess) rc = fail(NRQ_ERR_HIP, "cnn: event record failed");
    for (int r = 0; r < reps && !rc; ++r) rc = forward(n);
    float t = 0;
    if (!rc && (hipEventRecord(ev[1], stream()) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess ||
                hipEventElapsedTime(&t, ev[0], ev[1]) != hipSuccess))
        rc = fail(NRQ_ERR_HIP, "cnn: event timing failed");
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (rc) return rc;
    *ms = (double)t / reps;
    return NRQ_OK;
}

extern "C" int nrq_cnn_loss_grad(nrq_cnn *n, const float *in, const float *target, double norm, double *loss, float *grad)
{
    return nrq_cnn_loss_grad_trace(n, in, target, norm, loss, grad, nullptr, nullptr);
}

extern "C" int nrq_cnn_loss_grad_trace(nrq_cnn *n, const float *in, const float *target, double norm, double *loss,
                                       float *grad, float *acts, float *deltas)
{
    NEED_NET(n);
    if (!lat().initialised) return fail(NRQ_ERR_STATE, "lattice: not initialised");
    if (!in || !target || !loss || !grad || !(norm > 0)) return fail(NRQ_ERR_ARG, "cnn: bad argument");
    const geom g = lat().g;
    if (int e = ensure_geom(n, g)) return e;
    std::vector<bf16> a((size_t)g.vol * CH, 0);
    for (int64_t s = 0; s < g.vol; ++s)
        for (int c = 0; c < CIN; ++c) a[(size_t)s * CH + c] = f2bf(in[(size_t)s * CIN + c]);
    float *t = nullptr;
    NRQ_HIP(hipMalloc(&t, sizeof(float) * (size_t)g.vol * CIN));
    NRQ_HIP(hipMemcpyAsync(n->act[0], a.data(), a.size() * sizeof(bf16), hipMemcpyHostToDevice, stream()));
    NRQ_HIP(hipMemcpyAsync(t, target, sizeof(float) * (size_t)g.vol * CIN, hipMemcpyHostToDevice, stream()));
    NRQ_HIP(hipMemsetAsync(n->loss, 0, sizeof(double), stream()));
    const size_t lay = (size_t)g.vol * CH;
    std::vector<bf16> at(acts ? NL * lay : 0), dt(deltas ? NL * lay : 0);
    int e = forward(n);
    if (!e && acts)
        for (int l = 0; l < NL && !e; ++l)
            if (hipMemcpyAsync(at.data() + l * lay, n->act[l], sizeof(bf16) * lay, hipMemcpyDeviceToHost, stream()) !=
                hipSuccess)
                e = fail(NRQ_ERR_HIP, "cnn: activation copy failed");
    if (!e) e = backward(n, t, norm, deltas ? dt.data() : nullptr);
    if (!e) {
        NRQ_HIP(hipMemcpyAsync(grad, n->grad, sizeof(float) * NW, hipMemcpyDeviceToHost, stream()));
        NRQ_HIP(hipMemcpyAsync(loss, n->loss, sizeof(double), hipMemcpyDeviceToHost, stream()));
        NRQ_HIP(hipStreamSynchronize(stream()));
        for (size_t k = 0; k < at.size(); ++k) acts[k] = bf2f(at[k]);
        for (size_t k = 0; k < dt.size(); ++k) deltas[k] = bf2f(dt[k]);
    }
    (void)hipFree(t);
    return e;
}

extern "C" int nrq_cnn_harvest(nrq_cnn *n, const nrq_gauge *gf, double mass, double stepwidth, int nsolves, uint64_t seed)
{
    NEED_NET(n);
    if (!lat().initialised) return fail(NRQ_ERR_STATE, "lattice: not initialised");
    if (!gf || !gf->u[0]) return fail(NRQ_ERR_ARG, "cnn: null gauge configuration");
    if (!(mass > 0) || nsolves < 1 || !(stepwidth > 0)) return fail(NRQ_ERR_ARG, "cnn: bad mass, step width or solve count");
    const geom g = lat().g;
    if (g.n0 % 16) return fail(NRQ_ERR_UNSUPPORTED, "cnn: the x extent %d must be a multiple of 16", g.n0);
    if (!n->data.empty() && n->data_g.vol != g.vol) return fail(NRQ_ERR_STATE, "cnn: training set holds another lattice");
    n->data_g = g;
    constexpr int R = 12;
    const int64_t fs = 3 * g.vol;
    cplx<double> *psi = nullptr;
    NRQ_HIP(hipMalloc(&psi, sizeof(cplx<double>) * (size_t)fs * R));
    // CN coefficients of the quark (mass) and antiquark (-mass), c1 = 1
    const std::complex<double> aq(0.0, -(stepwidth * g.a[0] / 2.0) / mass), aa(0.0, (stepwidth * g.a[0] / 2.0) / mass);
    cplx<double> alpha[R], beta[R];
    cplx<double> *fields[R];
    for (int r = 0; r < R; ++r) {
        const std::complex<double> a = r < 6 ? aq : aa;
        alpha[r] = {a.real(), a.imag()};
        beta[r] = {a.real(), -a.imag()};
        fields[r] = psi + (int64_t)r * fs;
    }
    krylov_params p;
    p.kind = KRYLOV_GMRES_FP64;
    p.tol = 1e-12;
    int status = krylov_reserve(g.vol, p.restart, R);
    for (int k = 0; k < nsolves && !status; ++k) {
        hipLaunchKernelGGL(k_gauss, dim3(1024), dim3(256), 0, stream(), psi, fs * R, seed * 1000003ull + (uint64_t)k);
        nrq_solve_stats st = {0, 0, 0, 0.0};
        status = krylov_cn_step(gf, fields, R, alpha, beta, p, &st);
        if (status) break;
        // right-hand sides y of the solve (slot order), their features
        // beta K y and (beta K)^2 y, and the solutions in psi
        const cplx<double> *y = krylov_last_rhs();
        const cplx<double> *f1 = nullptr, *f2 = nullptr;
        status = krylov_rhs_features(gf, R, &f1, &f2);
        double yn2[R];
        if (!status) status = krylov_norms2(y, R, yn2);
        for (int q = 0; q < R && !status; ++q) {
            cnn_sample smp;
            if (hipMalloc(&smp.act, sizeof(bf16) * (size_t)g.vol * CH) != hipSuccess ||
                hipMalloc(&smp.tgt, sizeof(float) * (size_t)g.vol * CIN) != hipSuccess) {
                (void)hipGetLastError();
                if (smp.act) (void)hipFree(smp.act);
                status = fail(NRQ_ERR_ALLOC, "cnn: training set allocation failed");
                break;
            }
            const int64_t off = (int64_t)q * fs;
            hipLaunchKernelGGL(k_make_sample, dim3(1024), dim3(256), 0, stream(), y + off,
                               (const cplx<double> *)psi + off, f1 + off, f2 + off, 1.0 / std::sqrt(yn2[q]),
                               make_col_scale(beta[q], spectrum_top(g)), smp.act, smp.tgt, g.vol);
            n->data.push_back(smp);
        }
    }
    if (stream()) (void)hipStreamSynchronize(stream());
    (void)hipFree(psi);
    if (status) return status;
    NRQ_HIP(hipGetLastError());
    return NRQ_OK;
}

extern "C" int nrq_cnn_dataset_size(const nrq_cnn *n, int64_t *count)
{
    NEED_NET(n);
    if (!count) return fail(NRQ_ERR_ARG, "cnn: null argument");
    *count = (int64_t)n->data.size();
    return NRQ_OK;
}

extern "C" int nrq_cnn_train(nrq_cnn *n, int epochs, double lr, uint64_t seed, double *mean_loss)
{
    NEED_NET(n);
    if (n->data.empty()) return fail(NRQ_ERR_STATE, "cnn: empty training set");
    if (epochs < 1 || !(lr > 0)) return fail(NRQ_ERR_ARG, "cnn: bad epochs or learning rate");
    if (int e = ensure_geom(n, n->data_g)) return e;
    const geom &g = n->data_g;
    const size_t ns = n->data.size();
    std::vector<size_t> order(ns);
    for (size_t i = 0; i < ns; ++i) order[i] = i;
    uint64_t st = seed + 0x1234567ull;
    double last = 0;
    for (int ep = 0; ep < epochs; ++ep) {
        for (size_t i = ns - 1; i > 0; --i) { // Fisher-Yates with a splitmix stream
            st += 0x9e3779b97f4a7c15ull;
            uint64_t z = st;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            z ^= z >> 31;
            std::swap(order[i], order[(size_t)(z % (i + 1))]);
        }
        NRQ_HIP(hipMemsetAsync(n->loss, 0, sizeof(double), stream()));
        for (size_t i : order) {
            const cnn_sample &smp = n->data[i];
            NRQ_HIP(hipMemcpyAsync(n->act[0], smp.act, sizeof(bf16) * (size_t)g.vol * CH, hipMemcpyDeviceToDevice, stream()));
            if (int e = forward(n)) return e;
            if (int e = backward(n, smp.tgt, 1.0)) return e;
            ++n->adam_t;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.f / (1.f - std::pow(b1, (float)n->adam_t)), c2 = 1.f / (1.f - std::pow(b2, (float)n->adam_t));
            hipLaunchKernelGGL(k_adam, dim3((NW + 255) / 256), dim3(256), 0, stream(), n->w, (const float *)n->grad, n->m1,
                               n->m2, (float)lr, b1, b2, c1, c2, n->wb, n->wt);
            NRQ_HIP(hipGetLastError());
        }
        double tot = 0;
        NRQ_HIP(hipMemcpyAsync(&tot, n->loss, sizeof(double), hipMemcpyDeviceToHost, stream()));
        NRQ_HIP(hipStreamSynchronize(stream()));
        last = tot / (double)ns;
    }
    NRQ_HIP(hipMemcpyAsync(n->w_host.data(), n->w, sizeof(float) * NW, hipMemcpyDeviceToHost, stream()));
    NRQ_HIP(hipStreamSynchronize(stream()));
    if (mean_loss) *mean_loss = last;
    return NRQ_OK;
}

extern "C" int nrq_cn_set_preconditioner(nrq_cnn *n)
{
    g_precond = n;
    return NRQ_OK;
}
