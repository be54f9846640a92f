// nrq_cnn.h -- the learned preconditioner seen from the Krylov solver.
#pragma once
#include "nrq_internal.h"

struct nrq_cnn;

namespace nrq {

// z_r = v_r + N(v_r) for the R slot fields v (3V complex doubles each,
// colour major) at v[r], results at z[r]: the network acts on groups of three
// slots (18 real channels per site).  Needs n0 % 16 == 0.
int cnn_precondition(nrq_cnn *net, cplx<double> *const *v, cplx<double> *const *z, int R);

// the network registered with nrq_cn_set_preconditioner (may be null)
nrq_cnn *cn_preconditioner();

} // namespace nrq
