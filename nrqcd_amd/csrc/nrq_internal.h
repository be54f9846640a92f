// nrq_internal.h -- host-side state shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include "nrq_su3.h"
#include "../../include/nrqcd_amd.h"

namespace nrq {

// Error reporting: every entry point returns an NRQ_* status and leaves a
// message retrievable with nrq_last_error().
int fail(int code, const char *fmt, ...);
int check_hip(hipError_t e, const char *what);

#define NRQ_HIP(call)                                                     \
    do {                                                                  \
        hipError_t nrq_e_ = (call);                                       \
        if (nrq_e_ != hipSuccess) return ::nrq::check_hip(nrq_e_, #call); \
    } while (0)

struct lattice_state {
    bool initialised = false;
    geom g{};
};
lattice_state &lat();

// The library's single HIP stream (default priority); every kernel and copy
// is ordered on it.  Entry points that return host values synchronise it.
hipStream_t stream();
int ensure_device();

// Host ranlxd (level 1/2) used for the seeded initial condition.
int rng_init(int level, int64_t seed);
void rng_uniform(double *r, int64_t n);
bool rng_ready();

// Deterministic two-stage reductions: a fixed grid of REDUCE_BLOCKS blocks
// writes partials, one block folds them in a fixed order.
constexpr int REDUCE_BLOCKS = 1024;
constexpr int REDUCE_THREADS = 256;

// Per momentum index s: scale * sum_a Re[sum_ij F_i P_ij conj(F_j)] of the 24
// momentum-space fields fld[(a + 8 i) V + s] (device) with the real transverse
// projector of the backward-derivative momentum, into host out[V]
// (gaugeconfiguration_su3.f90:1954-1972).
int transverse_project(const cplx<double> *fld, double scale, double *out);

// One CN step of the live columns of n heavy fields with the full
// Hamiltonian (c2..c4 terms, nrqcd.f90:1175-1387): FP64 GMRES(restart) to
// true relative residual tol, at most maxit iterations per column.
int ham_cn_update(nrq_heavy *const *hs, const double *mass, int n, const nrq_gauge *gf, const double c[8],
                  double stepwidth, int restart, int64_t maxit, double tol, nrq_solve_stats *stats);

// Drops state that carries the lattice by value (captured leapfrog graphs);
// called by nrq_lattice_finalize.
void gauge_forget_lattice();

struct scratch_buf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int reserve(size_t n);
    ~scratch_buf();
};

} // namespace nrq

struct nrq_gauge {
    int precision = 64;       // 64: double storage, 32: float storage
    int64_t vol = 0;
    void *u[2] = {nullptr, nullptr}; // ping-pong link buffers [dir][elem][site] complex
    int cur = 0;              // which u[] holds the current links
    void *e = nullptr;        // E field [dir][a][site] real
    void *partials = nullptr; // reduction partials
    double *result = nullptr; // pinned host result slots
    int64_t steps_since_reunit = 0;
    int64_t reunit_interval = 0; // reunitarise after this many steps (0: never)
    double *dlog = nullptr;       // observables log on the device, rows of 5 raw sums
    int64_t log_cap = 0, log_n = 0; // rows allocated, rows held
};

struct nrq_heavy {
    int64_t vol = 0;
    void *psi = nullptr;      // 24 colour-vector fields [field][colour][site] complex double
    unsigned live = 0;        // bit f set: field f may be non-zero (clear bits are exactly zero)
    // Every change of psi increments version.  obs holds the lattice sums of
    // the correlator traces and the two squared norms of field version
    // obs_version, so the ten observables a driver reads per time step
    // (eight correlators, two norms) cost one pass over psi.
    uint64_t version = 0;
    uint64_t obs_version = ~(uint64_t)0;
    double obs[16] = {};
};
