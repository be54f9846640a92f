// nrq_krylov.h -- batched solvers for the CN systems of the NRQCD heavy
// fields, replacing PETSc KSP (nrqcd.f90:786-804, 1097-1161).
//
// K psi(x) = sum_i [psi(x) - (U_i(x) psi(x+i) + U_i(x-i)^+ psi(x-i))/2] / a_i^2
// is the matrix-free covariant Laplacian of the kinetic NRQCD Hamiltonian;
// each slot is a colour-vector field psi[slot][colour][site] with its own
// complex coefficients (quark and antiquark, or different masses, share one
// batched solve and one link stream).
#pragma once
#include "nrq_internal.h"
#include <vector>

namespace nrq {

constexpr int NFIELDS_PER_HEAVY = 24; // colour-vector fields per NRQCDField
constexpr int NSLOTS = 48;            // right-hand sides of one batched solve

enum krylov_kind {
    KRYLOV_GMRES_FP64 = 0,  // GMRES(m), classical Gram-Schmidt, FP64, zero guess
    KRYLOV_GMRES_MIXED = 1, // FP32 GMRES(m) corrections inside FP64 iterative refinement
    KRYLOV_CHEB_FP64 = 2,   // Chebyshev iteration on the known spectral segment, FP64
    KRYLOV_CHEB_MIXED = 3,  // FP32 Chebyshev corrections inside FP64 iterative refinement
    KRYLOV_FGMRES_CNN = 4,  // FGMRES(m) in FP64, right-preconditioned by the learned network
    KRYLOV_FGMRES_CNN_MIXED = 5, // FP32 FGMRES(m) with the network inside FP64 iterative refinement
};

struct krylov_params {
    int kind = KRYLOV_GMRES_FP64;
    int restart = 30;          // GMRES restart length m (PETSc default 30)
    int64_t maxit = 10000;     // cap on Krylov iterations per slot
    double tol = 1e-10;        // FP64 true relative residual
    double inner_tol = 1e-5;   // relative residual of each FP32 GMRES correction
    bool extrapolate = false;  // first-order initial guess instead of zero
};

// One CN step of R colour-vector fields (3V complex doubles each, colour
// major) at fields[q]: for q < R,
//   y_q = (1 + alpha_q K) psi_q,  solve (1 + beta_q K) x_q = y_q,  psi_q = x_q
// to an FP64 true relative residual <= p.tol.  Statistics accumulate into st.
// The workspace must hold R slots (krylov_reserve).
int krylov_cn_step(const nrq_gauge *gf, cplx<double> *const *fields, int R, const cplx<double> *alpha,
                   const cplx<double> *beta, const krylov_params &p, nrq_solve_stats *st);

// workspace for lattice volume vol, GMRES restart m and at least `slots`
// right-hand sides (0: per-slot scalars only)
int krylov_reserve(int64_t vol, int m, int slots);
void krylov_release();
// right-hand sides y of the last krylov_cn_step, slot-major (3V complex each)
const cplx<double> *krylov_last_rhs();
// squared norms of nf consecutive FP64 fields into out[nf] (host)
int krylov_norms2(const cplx<double> *x, int nf, double *out);
// Features of the right-hand sides y of the last krylov_cn_step on gf:
// f1_r = beta_r K y_r and f2_r = beta_r K f1_r (FP64, slot-major like y),
// valid until the next solve.
int krylov_rhs_features(const nrq_gauge *gf, int R, const cplx<double> **f1, const cplx<double> **f2);
// partial sums for the correlator kernel: device scratch of
// KRYLOV_REDUCE_DOUBLES doubles
constexpr size_t KRYLOV_REDUCE_DOUBLES = (size_t)NSLOTS * 256 * 8;
double *krylov_reduce_scratch();
double *krylov_host_scratch(); // pinned, >= 64 doubles

} // namespace nrq
