// nrq_grouplog.h -- principal logarithm of a unitary 3x3 matrix on the GPU
// (ArgLogU, matrixoperations.f90:474-495: S^+ diag(atan2(Im, Re) of the
// eigenvalues) S), used by the transverse correlators (GetGroupLog,
// su3.f90:318-334) and by the field strengths of the NRQCD Hamiltonian
// (gaugeconfiguration_su3.f90:1563-1581).
//
// U is normal, so its eigenvectors are those of the Hermitian matrix
// M = (U - U^+)/(2i) + gamma (U + U^+)/2, whose eigenvalue for eigenphase t is
// sin t + gamma cos t.  M is diagonalised by cyclic complex Jacobi rotations;
// each eigenvector's Rayleigh quotient with U gives the eigenphase.  Two
// eigenphases with t_j + t_k = pi - 2 atan(gamma) share an M eigenvalue, so a
// second gamma is tried when the eigen-residual of the first is large.
#pragma once
#include "nrq_su3.h"

namespace nrq {

__device__ __forceinline__ double nrm2c(cplx<double> a) { return a.re * a.re + a.im * a.im; }

__device__ inline void jacobi_herm3(mat3<double> &A, mat3<double> &V)
{
    set_zero(V);
    V.m[0] = V.m[4] = V.m[8] = {1.0, 0.0};
    const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
    for (int sweep = 0; sweep < 20; ++sweep) {
        const double off = nrm2c(A.m[1]) + nrm2c(A.m[2]) + nrm2c(A.m[5]);
        const double dia = A.m[0].re * A.m[0].re + A.m[4].re * A.m[4].re + A.m[8].re * A.m[8].re;
        if (!(off > 1e-34 * dia) || off == 0.0) break;
        for (int k = 0; k < 3; ++k) {
            const int p = P[k], q = Q[k];
            const cplx<double> b = A.m[3 * p + q];
            const double ab = sqrt(nrm2c(b));
            if (ab == 0.0) continue;
            const cplx<double> e = {b.re / ab, b.im / ab};
            const cplx<double> ec = {e.re, -e.im};
            const double th = 0.5 * atan2(2.0 * ab, A.m[4 * p].re - A.m[4 * q].re);
            double s, c;
            sincos(th, &s, &c);
            // columns: X <- X J   (X = A, then V)
            for (int r = 0; r < 3; ++r) {
                const cplx<double> xp = A.m[3 * r + p], xq = A.m[3 * r + q];
                A.m[3 * r + p] = c * xp + s * (ec * xq);
                A.m[3 * r + q] = (-s) * xp + c * (ec * xq);
                const cplx<double> vp = V.m[3 * r + p], vq = V.m[3 * r + q];
                V.m[3 * r + p] = c * vp + s * (ec * vq);
                V.m[3 * r + q] = (-s) * vp + c * (ec * vq);
            }
            // rows: A <- J^+ A
            for (int col = 0; col < 3; ++col) {
                const cplx<double> ap = A.m[3 * p + col], aq = A.m[3 * q + col];
                A.m[3 * p + col] = c * ap + s * (e * aq);
                A.m[3 * q + col] = (-s) * ap + c * (e * aq);
            }
            A.m[3 * p + q] = A.m[3 * q + p] = {0.0, 0.0};
            A.m[4 * p].im = A.m[4 * q].im = 0.0;
        }
    }
}

// H = -i log U on the principal branch (eigenphases in (-pi, pi]), the full
// Hermitian matrix including its trace.
__device__ inline mat3<double> group_log_matrix(const mat3<double> &U)
{
    const double gammas[2] = {0.41421356237309503, -0.7};
    mat3<double> best_v;
    cplx<double> best_l[3];
    double best_res = 1e300;
    for (int t = 0; t < 2 && best_res > 1e-12; ++t) {
        const double gm = gammas[t];
        mat3<double> M;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const cplx<double> a = U.m[3 * r + c], b = U.m[3 * c + r]; // b^* = (U^+)_rc
                // (U - U^+)/(2i) + gm (U + U^+)/2
                const cplx<double> k = {0.5 * (a.im + b.im), 0.5 * (-a.re + b.re)};
                const cplx<double> h = {0.5 * (a.re + b.re), 0.5 * (a.im - b.im)};
                M.m[3 * r + c] = k + gm * h;
            }
        mat3<double> V;
        jacobi_herm3(M, V);
        double res = 0;
        cplx<double> lam[3];
        for (int k = 0; k < 3; ++k) {
            cplx<double> uv[3];
            for (int r = 0; r < 3; ++r) {
                cplx<double> acc = {0.0, 0.0};
                for (int c = 0; c < 3; ++c) mac(acc, U.m[3 * r + c], V.m[3 * c + k]);
                uv[r] = acc;
            }
            cplx<double> l = {0.0, 0.0};
            for (int r = 0; r < 3; ++r) mac_cn(l, V.m[3 * r + k], uv[r]); // v^+ U v
            lam[k] = l;
            double rk = 0;
            for (int r = 0; r < 3; ++r) rk += nrm2c(uv[r] - l * V.m[3 * r + k]);
            res = fmax(res, rk);
        }
        if (res < best_res) {
            best_res = res;
            best_v = V;
            for (int k = 0; k < 3; ++k) best_l[k] = lam[k];
        }
    }
    mat3<double> H;
    set_zero(H);
    for (int k = 0; k < 3; ++k) {
        const double th = atan2(best_l[k].im, best_l[k].re);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                cplx<double> vv = {0.0, 0.0};
                mac_nc(vv, best_v.m[3 * r + k], best_v.m[3 * c + k]); // v_r conj(v_c)
                H.m[3 * r + c] = H.m[3 * r + c] + th * vv;
            }
    }
    return H;
}

// Algebra coordinates 2 Re Tr(T^a H) of H = -i log U (GetGroupLog).
__device__ inline void group_log_coords(const mat3<double> &U, double *out)
{
    two_re_trace_gen(group_log_matrix(U), out);
}

} // namespace nrq
