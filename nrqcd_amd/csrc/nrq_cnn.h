// nrq_cnn.h -- the learned preconditioner seen from the Krylov solver.
#pragma once
#include "nrq_internal.h"

struct nrq_cnn;

namespace nrq {

// z_r = v_r - f1_r + f2_r + N(v_r, f1_r, f2_r) for the slots r < R with
// act[r] (all when act is null): v_r, f1_r = beta_r K v_r, f2_r = beta_r K f1_r
// are colour fields of 3V complex numbers (colour major), z_r receives the
// result.  v - f1 + f2 is the second-order Neumann polynomial of
// (1 + beta K)^-1; the network corrects it in mass-independent units (its
// inputs v, K v / kmax, K^2 v / kmax^2 and its output times -(beta kmax)^3,
// beta = beta[r] on the host).  The network takes one colour field with its
// two features per pass (18 real channels per site).  Needs n0 % 16 == 0.
template <typename T>
int cnn_precondition(nrq_cnn *net, const cplx<T> *const *v, const cplx<T> *const *f1, const cplx<T> *const *f2,
                     cplx<T> *const *z, int R, const int *act, const cplx<double> *beta);

// the network registered with nrq_cn_set_preconditioner (may be null)
nrq_cnn *cn_preconditioner();

} // namespace nrq
