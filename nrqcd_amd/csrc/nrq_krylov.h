// nrq_krylov.h -- batched Krylov solver for the CN systems (1 + b_r K) x_r = y_r
// of R right-hand-side slots, replacing PETSc KSP (nrqcd.f90:786-804).
//
// K psi(x) = sum_i [psi(x) - (U_i(x) psi(x+i) + U_i(x-i)^+ psi(x-i))/2] / a_i^2
// is the matrix-free covariant Laplacian of the kinetic NRQCD Hamiltonian;
// each slot is a colour-vector field psi[slot][colour][site] and has its own
// complex coefficient b_r (quark and antiquark, or different masses, share one
// batched solve and one link stream).
#pragma once
#include "nrq_internal.h"
#include <vector>

namespace nrq {

constexpr int NSLOTS = 24; // colour-vector fields per NRQCDField (and max batch)

enum krylov_kind {
    KRYLOV_GMRES_FP64 = 0,  // GMRES(m), classical Gram-Schmidt, FP64, zero guess
    KRYLOV_GMRES_MIXED = 1, // FP32 GMRES(m) corrections inside FP64 iterative refinement
    KRYLOV_CHEB_FP64 = 2,   // Chebyshev iteration on the known spectral segment, FP64
    KRYLOV_CHEB_MIXED = 3,  // FP32 Chebyshev corrections inside FP64 iterative refinement
};

struct krylov_params {
    int kind = KRYLOV_GMRES_FP64;
    int restart = 30;          // GMRES restart length m (PETSc default 30)
    int64_t maxit = 10000;     // cap on Krylov iterations per slot
    double tol = 1e-10;        // FP64 true relative residual
    double inner_tol = 1e-5;   // relative residual of each FP32 correction solve
    bool extrapolate = false;  // first-order initial guess instead of zero
};

// Solves (1 + beta_r K) x_r = y_r for r < R.  y (slot-major, 3V complex
// doubles each) must be in krylov_rhs(); the solution is left in
// krylov_solution().  `prev` (may be null) holds the fields whose CN update
// produced y (used by the extrapolated initial guess with coefficient
// alpha_r: y = (1 + alpha_r K) prev).
int krylov_solve(const nrq_gauge *gf, int R, const cplx<double> *beta, const cplx<double> *alpha,
                 const cplx<double> *prev_slots, const double *y_norm2, const krylov_params &p, nrq_solve_stats *st);

int krylov_reserve(int64_t vol, int m);
void krylov_release();
cplx<double> *krylov_rhs();
cplx<double> *krylov_solution();
cplx<double> *krylov_scratch(); // NSLOTS x 3V complex doubles

// y_r = x_r + beta_r K x_r for r < nf (FP64 fields); beta on the host.
int krylov_apply(const nrq_gauge *gf, const cplx<double> *x, cplx<double> *y, int nf, const cplx<double> *beta);
// squared norms of nf FP64 fields into out[nf] (host)
int krylov_norms2(const cplx<double> *x, int nf, double *out);
// partial sums for the correlator kernel: device scratch of >= 8 x 256 doubles
double *krylov_reduce_scratch();
double *krylov_host_scratch(); // pinned, >= 64 doubles

} // namespace nrq
