/*
 * nrqcd_amd.h -- C ABI of the MI355X-native real-time SU(3) leapfrog and
 * NRQCD Crank-Nicolson stepper.
 *
 * This is the drop-in boundary for the reference's Fortran module interfaces
 * `lattice`, `random`, `gaugeconfiguration_su3` and `nrqcd`: the Fortran
 * modules under nrqcd_amd/fortran/ keep the reference's module, type and
 * type-bound-procedure names and bind these functions with ISO_C_BINDING.
 * Each declaration cites the reference interface it replaces.
 *
 * Conventions
 *  - Every function returns NRQ_OK (0) or a negative NRQ_ERR_* code; the
 *    message of the last failure is available from nrq_last_error().
 *  - Lattice indices are 1-based int64 with x fastest (lattice.f90:883-931);
 *    directions are +-1..3 (gaugeconfiguration_su3.f90:193-199).
 *  - Complex numbers are (re, im) pairs of doubles.  Host arrays exchanged
 *    with the library use the reference's Fortran layouts:
 *      links  (3,3,3,V) complex128  gaugeconfiguration_su3.f90:35,134-137
 *      efield (8,3,V)   real64
 *      prop   (6,6,V)   complex128  nrqcd.f90:60-63
 *  - Device state stays resident in HBM between calls.  All work is ordered
 *    on one HIP stream; calls that return host values synchronise it.
 *  - One host thread; handles are not thread-safe.
 */
#ifndef NRQCD_AMD_H
#define NRQCD_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NRQ_OK 0
#define NRQ_ERR_ARG (-1)          /* invalid argument */
#define NRQ_ERR_STATE (-2)        /* module not initialised / already initialised */
#define NRQ_ERR_HIP (-3)          /* HIP runtime or hipFFT failure */
#define NRQ_ERR_UNSUPPORTED (-4)  /* unsupported request (unknown solver kind, CNN on an x extent not a multiple of 16) */
#define NRQ_ERR_NOCONV (-5)       /* Krylov solver hit maxit before the tolerance */
#define NRQ_ERR_ALLOC (-6)        /* device or host allocation failed */

#define NRQ_PRECISION_FP64 64
#define NRQ_PRECISION_FP32 32

/* Copy the last error message (NUL-terminated, truncated to n). Returns the
 * code of the last failure (0 if none). */
int nrq_last_error(char *buf, size_t n);

/* Library version string. */
const char *nrq_version(void);

/* Timing and ordering on the library's HIP stream (no reference
 * counterpart; the reference times nothing).  nrq_event_record records HIP
 * event `slot` (0..15) on the stream; nrq_event_elapsed waits for `stop` and
 * returns the milliseconds between two recorded slots; nrq_synchronize waits
 * for all queued work. */
int nrq_event_record(int slot);
int nrq_event_elapsed(int start, int stop, double *ms);
int nrq_synchronize(void);
/* Runs a plain streaming-copy kernel (read + write `bytes`) `reps` times and
 * returns the achieved read+write bandwidth in GB/s: the practical HBM
 * ceiling of this device and a known-size reference for hardware byte
 * counters.  No reference counterpart. */
int nrq_bandwidth_probe(int64_t bytes, int reps, double *gbs);

/* ---- lattice (lattice.f90) ----------------------------------------- */

/* lattice::InitModule(LatticeExtensions_(3) int64, LatticeSpacings_(0:3))
 * lattice.f90:124-228.  spacing = (a0, a1, a2, a3). Single process: the
 * partition and halo of lattice.f90:151-192 do not exist. */
int nrq_lattice_init(const int64_t ext[3], const double spacing[4]);
/* Releases the module state (no reference counterpart; lets a process
 * re-initialise with another lattice). */
int nrq_lattice_finalize(void);
/* lattice::GetLatticeSize, lattice.f90 (product of extents) */
int nrq_lattice_size(int64_t *vol);
/* lattice::GetLatticeIndex, lattice.f90:476-484 */
int nrq_lattice_index(const int64_t pos[3], int64_t *idx);
/* lattice::GetLatticePosition, lattice.f90:794-804 */
int nrq_lattice_position(int64_t idx, int64_t pos[3]);
/* lattice::GetNeib_G, lattice.f90:531-543 (dir in -3..3) */
int nrq_lattice_neib(int dir, int64_t idx, int64_t *out);
/* lattice::GetNegativeLatticeIndex, lattice.f90:495-505 */
int nrq_lattice_negative_index(int64_t idx, int64_t *out);
/* lattice::GetMomentum_G (backward derivative), lattice.f90:963-979;
 * p = 3 complex numbers */
int nrq_lattice_momentum(int64_t idx, double p[6]);
/* lattice::GetNorm2Momentum_G, lattice.f90:1048-1058 */
int nrq_lattice_norm2_momentum(int64_t idx, double *out);
/* lattice::GetMaxNorm2Momentum, lattice.f90:1101-1105 (largest |p| above) */
int nrq_lattice_max_norm2_momentum(double *out);

/* ---- random (random.f90 / ranlxd_generator.f90) --------------------- */

/* random::InitModule(CustomSeed) -> ranlxd_init(2, seed), random.f90:126-140 */
int nrq_random_init(int64_t seed);
/* random::GetRandomUniformReal(n) -> ranlxd(r), random.f90:148-175 */
int nrq_random_uniform(double *r, int64_t n);

/* ---- gaugeconfiguration_su3 ------------------------------------------ */

typedef struct nrq_gauge nrq_gauge;

/* GaugeConfiguration%Allocate, gaugeconfiguration_su3.f90:128-138.
 * precision: NRQ_PRECISION_FP64 or NRQ_PRECISION_FP32 (link/E storage). */
int nrq_gauge_create(int precision, nrq_gauge **h);
/* GaugeConfiguration%Deallocate, gaugeconfiguration_su3.f90:145-151 */
int nrq_gauge_destroy(nrq_gauge *h);
/* Intrinsic assignment GaugeConf_t2 = GaugeConf_t1 (main.f90:2320): deep
 * device-to-device copy; dst is re-allocated to src's precision. */
int nrq_gauge_copy(nrq_gauge *dst, const nrq_gauge *src);
/* GaugeConfiguration%ColdInit, gaugeconfiguration_su3.f90:363-378 */
int nrq_gauge_cold(nrq_gauge *h);
/* GaugeConfiguration%TransversePolarisedOccupiedInit_Box(SaturationScale,
 * Amplitude, Coupling), gaugeconfiguration_su3.f90:906-951 and 1263-1446.
 * Draws from the library's ranlxd stream (nrq_random_init). */
int nrq_gauge_box_ic(nrq_gauge *h, double saturation_scale, double amplitude, double coupling);
/* The same with its optional outputs aa_correlator, ee_correlator
 * (gaugeconfiguration_su3.f90:906-951 -> 1390-1421): per momentum index
 * (lattice index - 1) sum_a Re[X_i P_ij conj(X_j)] / V / 8 of the
 * momentum-space fields X = A, E with the real transverse projector P and the
 * physical volume V; either pointer may be NULL. */
int nrq_gauge_box_ic_correlators(nrq_gauge *h, double saturation_scale, double amplitude, double coupling, double *aa,
                                 double *ee);
/* TransversePolarisedOccupiedInit(Occupation, [aa], [ee]),
 * gaugeconfiguration_su3.f90:1263-1446 (specific_sender = 0), or
 * TransversePolarisedOccupiedInit_specificSender, :959-1255 (specific_sender
 * = 1: complex amplitudes drawn in momentum space in lattice order, then made
 * Hermitian), for any occupation[V] >= 0 by lattice index.  Both are private
 * in the reference, which reaches the first through the box form.  Fails
 * when the random module is not initialised, as the reference does. */
int nrq_gauge_transverse_init(nrq_gauge *h, const double *occupation, int specific_sender, double *aa, double *ee);
/* GaugeConfiguration%EquilibriumInit(Beta, nefieldinit, nequilibrium,
 * [tolerance_Eprojection], [ChargeDensity], ...), gaugeconfiguration_su3.f90:
 * 380-814: from cold links, nefieldinit times thermal E (sigma = sqrt(4/beta)),
 * Gauss-law projection to max 2|Tr(T^a G)| <= tolerance (<= 0: the
 * reference's default 1E-3 as a single-precision literal), nequilibrium
 * leapfrog steps.  charge_density: NULL or 3 x V (colour fastest, the
 * reference's ChargeDensity(colour, index)).  Optional per-iteration outputs
 * (nefieldinit entries, NULL to skip): energy and Gauss deviation after the
 * leapfrog steps, projection sweeps.  The projection is capped at 10^6 sweeps
 * (NRQ_ERR_NOCONV); the reference has no cap. */
int nrq_gauge_equilibrium_init(nrq_gauge *h, double beta, int64_t nefieldinit, int64_t nequilibrium, double tolerance,
                               const double *charge_density, double *energies, double *deviations,
                               int64_t *projections);
/* Upload / download in the reference layouts links(3,3,3,V) complex128 and
 * efield(8,3,V) real64 (gaugeconfiguration_su3.f90:35-37).  Either pointer may
 * be NULL to skip that field. */
int nrq_gauge_upload(nrq_gauge *h, const double *links, const double *efield);
int nrq_gauge_download(const nrq_gauge *h, double *links, double *efield);
/* GaugeConfiguration%GetLink_G(i, LatticeIndex), gaugeconfiguration_su3.f90:
 * 160-196: dir 0 -> identity, dir < 0 -> U_{-dir}(x-dir)^dagger.
 * link = 3x3 complex, column-major. */
int nrq_gauge_get_link(const nrq_gauge *h, int dir, int64_t idx, double link[18]);
/* GetEfield_G(a, i, LatticeIndex), gaugeconfiguration_su3.f90:239-254;
 * e = 8 reals */
int nrq_gauge_get_efield(const nrq_gauge *h, int dir, int64_t idx, double e[8]);
/* GaugeConfiguration%Update([StepWidth]) -> Update_Leapfrog applied nsteps
 * times, gaugeconfiguration_su3.f90:2082-2111 (links then E for
 * stepwidth > 0, E then links otherwise). */
int nrq_gauge_leapfrog(nrq_gauge *h, double stepwidth, int64_t nsteps);
/* nrq_gauge_leapfrog for n configurations of one storage precision on the
 * current lattice (an ensemble): each pass over all of them is one launch,
 * and the launch sequence of nsteps steps (up to 256) is captured into a
 * hipGraph and replayed on later calls with the same buffers (NRQ_GRAPH=0 in
 * the environment launches directly, as do longer sequences).  Each configuration's reunitarisation
 * interval is honoured; when all members are at the same point of the same
 * interval, forward steps project inside the fused passes as
 * nrq_gauge_leapfrog does.  No reference counterpart: the reference evolves
 * one configuration per run (src/main.f90:2572-2620 loops over runs). */
int nrq_gauge_leapfrog_batch(nrq_gauge *const *h, int n, double stepwidth, int64_t nsteps);
/* Kernel timer of the fused leapfrog pass (measurement only, no reference
 * counterpart).  After nrq_leapfrog_timer_start(max_spans), each leapfrog
 * call that is not graph-captured records a HIP event pair on the library
 * stream around its run of fused launches (the E update of one step fused
 * with the link update of the next), up to max_spans runs.
 * nrq_leapfrog_timer_stop synchronises and returns the summed event time,
 * the fused launches and (site, configuration) steps those runs hold, and
 * the fused launches of runs past the last span (not timed). */
int nrq_leapfrog_timer_start(int64_t max_spans);
int nrq_leapfrog_timer_stop(double *fused_ms, int64_t *launches, int64_t *site_steps, int64_t *skipped);
/* GaugeConfiguration%GetEnergy(), gaugeconfiguration_su3.f90:1769-1817 */
int nrq_gauge_energy(const nrq_gauge *h, double *energy);
/* GaugeConfiguration%GetDeviationFromGaussLaw(), gaugeconfiguration_su3.f90:
 * 1480-1531 (sum of 2|Im Tr(T^a D)|; zero up to rounding by construction) */
int nrq_gauge_gauss_ref(const nrq_gauge *h, double *gauss);
/* Gauss-law constraint max_{x,a} 2|Tr(T^a G(x))| of GetG/GetGdeviation,
 * gaugeconfiguration_su3.f90:724-813 (no charge density) */
int nrq_gauge_gauss_max(const nrq_gauge *h, double *gauss);
/* (1/3V) sum_{x,i<j} Re Tr P_ij(x)/3 built from GetPlaquette_M,
 * gaugeconfiguration_su3.f90:1722-1737 */
int nrq_gauge_mean_plaquette(const nrq_gauge *h, double *plaq);
/* All four observables in one pass: out = {energy, gauss_ref, gauss_max,
 * mean_plaquette}. */
int nrq_gauge_observables(const nrq_gauge *h, double out[4]);
/* Observables log kept on the device.  nrq_gauge_log_observables enqueues
 * the computation of the four observables above for the current
 * configuration and appends them to the handle's log without waiting for the
 * GPU, so a driver that logs every n steps (the reference's drivers compute
 * and write energy, Gauss law and plaquette at every logged step,
 * src/main.f90:2609-2620 with gaugeconfiguration_su3.f90:1480-1531,
 * 1722-1737, 1769-1817) keeps the leapfrog queue full.  nrq_gauge_log_read
 * waits for the GPU and moves the oldest min(held, max_rows) rows to out (4
 * doubles per row, as nrq_gauge_observables), sets *rows and drops them from
 * the log.  The log grows as needed (device memory, 40 B per row) and is
 * freed with the handle.  nrq_gauge_copy does not copy it; a copy that has to
 * reallocate its destination (other precision or lattice) clears the
 * destination's log.  No reference counterpart beyond the quantities
 * themselves. */
int nrq_gauge_log_observables(nrq_gauge *h);
int nrq_gauge_log_read(nrq_gauge *h, double *out, int64_t max_rows, int64_t *rows);
/* GaugeConfiguration%GetPlaquette_G(i, j, LatticeIndex) for spatial i, j,
 * gaugeconfiguration_su3.f90:1591-1615 (3x3 complex column-major) */
int nrq_gauge_plaquette(const nrq_gauge *h, int i, int j, int64_t idx, double P[18]);
/* Projects every link back onto SU(3) in FP64 (Gram-Schmidt of rows 1-2,
 * row 3 = conj(row1 x row2)).  No reference counterpart: the reference keeps
 * FP64 links and never reunitarises; used by the FP32 storage mode. */
int nrq_gauge_reunitarise(nrq_gauge *h);
/* Makes nrq_gauge_leapfrog reunitarise (as above) after every `interval`
 * steps, counted across calls (0: never, the default).  In a forward call a
 * point inside the call is done by the fused pass that follows it (the links
 * are projected after the E update has used them and before the next link
 * update, the same result as a separate pass there); a point at the call's
 * end by a separate pass.  No reference counterpart: it keeps FP32 link
 * storage on the group over long runs. */
int nrq_gauge_set_reunitarise_interval(nrq_gauge *h, int64_t interval);
/* Launch shape the leapfrog autotuner chose for this configuration's
 * precision and lattice size: threads per block, XCD remap factor (1 = none),
 * tiling (0: consecutive sites, 1: z-sweep tiles of full x-rows, 2: z-sweep
 * tiles 16 sites wide, 3 and 4: LDS-staged 16 x 8 tiles over runs of 16 or 8
 * z planes, three planes of links in LDS, FP32 storage only, extents
 * multiples of 16, 8 and the run), streaming
 * (1: nontemporal loads/stores of E and the
 * new links) and two_row (1: neighbour links read as two rows, the third
 * rebuilt as the conjugate cross product, exact for SU(3)).  The first
 * leapfrog call per (precision, size) times every shape on scratch copies;
 * NRQ_AUTOTUNE=0 in the environment disables it and fixes the default shape
 * (256 threads, consecutive sites, plain loads), NRQ_LF_SHAPE=n pins shape n.
 * All shapes compute the same update, equal up to rounding.  No reference
 * counterpart. */
int nrq_gauge_launch_shape(const nrq_gauge *h, int *block_sites, int *xcd_remap, int *tiling, int *streaming,
                           int *two_row);
/* The autotuner's candidate shapes: their number, the parameters of shape k
 * (as nrq_gauge_launch_shape) and whether it applies to the current lattice
 * (bit 0: with FP64 storage, bit 1: with FP32 storage).
 * The tuner times only the shapes whose one fused pass agrees with the
 * default shape's (1e-11 FP64, 5e-5 FP32; two-row shapes rebuild a row and
 * agree to rounding) and reports any other on stderr.  No reference
 * counterpart. */
int nrq_leapfrog_shape_count(int *n);
int nrq_leapfrog_shape_info(int k, int *block_sites, int *xcd_remap, int *tiling, int *streaming, int *two_row,
                            int *applies);
/* Pins launch shape k (0 .. count-1) for every later leapfrog call, or k = -1
 * for the autotuner's choice (the default; NRQ_LF_SHAPE=k in the environment
 * sets the initial value).  A leapfrog call fails with NRQ_ERR_UNSUPPORTED
 * while a pinned shape does not apply to the lattice.  No reference
 * counterpart. */
int nrq_leapfrog_set_shape(int k);
int nrq_leapfrog_get_shape(int *k);
/* GaugeConfiguration%CoulombGaugefixing(Tolerance_, alpha_, MaxIterations_),
 * gaugeconfiguration_su3.f90:2269-2497 (the reference's defaults are 1e-8,
 * 0.08 and 100): Davies' Fourier-accelerated gauge fixing of the links.  E is
 * not transformed, as in the reference.  Stops when max_x ||Delta(x)||_F / 9
 * <= tol or after max_iterations gauge transformations; iterations (may be
 * null) receives the transformations applied, final_deviation (may be null)
 * the last max_x ||Delta(x)||_F / 9. */
int nrq_gauge_coulomb_gaugefix(nrq_gauge *h, double tol, double alpha, int64_t max_iterations, int64_t *iterations,
                               double *final_deviation);
/* GaugeConfiguration%GetTransverseAACorrelator / GetTransverseEECorrelator,
 * gaugeconfiguration_su3.f90:1825-1972: V doubles each, indexed by momentum
 * index (lattice index - 1); either pointer may be null.  The driver's gluon
 * distribution is sqrt(aa ee) per momentum next to nrq_lattice_norm2_momentum
 * (main.f90:3093-3189). */
int nrq_gauge_transverse_correlators(const nrq_gauge *h, double *aa, double *ee);

/* ---- nrqcd ------------------------------------------------------------ */

typedef struct nrq_heavy nrq_heavy;

typedef struct {
    int64_t iterations;      /* Krylov iterations summed over the solved columns */
    int64_t solves;          /* columns with a non-zero right-hand side */
    int64_t restarts;        /* GMRES restart cycles (incl. refinement passes) */
    double max_true_relres;  /* worst FP64 true relative residual ||b-Ax||/||b|| */
} nrq_solve_stats;

#define NRQ_SOLVER_GMRES_FP64 0   /* GMRES(m), classical Gram-Schmidt, FP64, zero guess */
#define NRQ_SOLVER_GMRES_MIXED 1  /* FP32 GMRES(m) corrections inside FP64 iterative
                                     refinement on the true residual, first-order
                                     initial guess */
#define NRQ_SOLVER_CHEBYSHEV_FP64 2  /* Chebyshev iteration on the known spectral
                                        segment of 1 + i dt/2 H (no inner products),
                                        FP64, first-order guess, GMRES fallback */
#define NRQ_SOLVER_CHEBYSHEV_MIXED 3 /* FP32 Chebyshev corrections inside FP64
                                        iterative refinement */
#define NRQ_SOLVER_FGMRES_CNN 4      /* FGMRES(m) in FP64, right-preconditioned by
                                        the learned network (nrq_cn_set_preconditioner) */
#define NRQ_SOLVER_FGMRES_CNN_MIXED 5 /* FP32 FGMRES(m) corrections with the learned
                                         network inside FP64 iterative refinement */

/* nrqcd::InitModule(Mass, WilsonCoeffs(4), StepWidth, kspTol),
 * nrqcd.f90:167-192 with KSPSetTolerances(rtol=kspTol), nrqcd.f90:803-804.
 * c = 4 complex Wilson coefficients.  The solver judges convergence on the
 * FP64 true relative residual (stricter than PETSc's preconditioned test). */
int nrq_cn_init(double mass, const double c[8], double stepwidth, double ksp_tol);
/* nrqcd::FinalizeModule, nrqcd.f90:199-212 */
int nrq_cn_finalize(void);
/* Solver selection: kind NRQ_SOLVER_*, GMRES restart length (PETSc default
 * 30) and iteration cap (PETSc default maxits 1e4, nrqcd.f90:803-804). */
int nrq_cn_set_solver(int kind, int restart, int64_t maxit);
/* Launch shape of the CN stencil kernels: colour-vector fields per thread
 * sharing each link load in FP64 (nc64) and FP32 (nc32) arithmetic (1, 2 or
 * 4), the site-block tiling (0: consecutive sites, 1: z-sweep tiles of
 * x-rows, 2: z-sweep 16-wide tiles, 3: wave items, 64 consecutive sites per
 * block with one field chunk per wavefront, so the chunks of a site share
 * their link loads in the CU's L1; 4: wave items over 64-site blocks of
 * band-planes, runs of up to 8 x-rows at one z, ordered by band and z, so
 * that with the XCD remap each XCD sweeps whole bands through z) and the XCD
 * remap (1 or 8).  Every shape
 * computes the same per-field sums in the same order.  NRQ_CN_SHAPE=
 * "nc64,nc32,tiling,xcd" in the environment sets the initial value (default
 * 2,4,4,8).  No reference counterpart. */
int nrq_cn_set_shape(int nc64, int nc32, int tiling, int xcd_remap);
int nrq_cn_get_shape(int *nc64, int *nc32, int *tiling, int *xcd_remap);
/* Column solves that NRQ_SOLVER_FGMRES_CNN handed to plain GMRES since the
 * library was loaded: the network made the Arnoldi step non-finite, a
 * restart cycle reduced the residual less than tenfold, or four cycles did
 * not converge.  The plain solve continues from the iterate reached, so the
 * answer is the same; only the work changes.  (PETSc's Jacobi preconditioner
 * scales by the constant diagonal of this operator, which leaves GMRES
 * unchanged.)  No reference counterpart (nrqcd.f90:786-804 configures
 * PETSc's preconditioner). */
int nrq_cn_fallbacks(int64_t *count);

/* NRQCDField() constructor / Destructor, nrqcd.f90:234-260 */
int nrq_heavy_create(nrq_heavy **h);
int nrq_heavy_destroy(nrq_heavy *h);
/* NRQCDField assignment(=) -> assign, nrqcd.f90:219-228 */
int nrq_heavy_copy(nrq_heavy *dst, const nrq_heavy *src);
/* NRQCDField%InitSinglePoint(LatticeIndex_quark, LatticeIndex_antiq),
 * nrqcd.f90:324-350: 1_6/sqrt(6) at the given sites */
int nrq_heavy_init_point(nrq_heavy *h, int64_t idx_quark, int64_t idx_antiq);
/* NRQCDField%InitSinglePointSingleDoF, nrqcd.f90:357-393 (spin 1..2,
 * colour 1..3, spin-colour index colour + 3(spin-1) as nrqcd.f90:437-444) */
int nrq_heavy_init_point_dof(nrq_heavy *h, int spin_q, int colour_q, int64_t idx_q,
                             int spin_a, int colour_a, int64_t idx_a);
/* Direct access to the public QuarkProp / AntiQprop components
 * (nrqcd.f90:60-63) in the (6,6,V) complex128 layout; NULL skips. */
int nrq_heavy_upload(nrq_heavy *h, const double *quark, const double *antiq);
int nrq_heavy_download(const nrq_heavy *h, double *quark, double *antiq);
/* NRQCDField%Update(GaugeConf, Mass, WilsonCoeffs, [StepWidth]) ->
 * Update_CrankNicholson, nrqcd.f90:1053-1168: rhs = S psi, solve
 * S^dagger psi' = rhs with S = 1 - i (StepWidth a0/2) H, H = -c1 D^2/2M; the
 * antiquark uses -M, conj(c).  With c = (c1, 0, 0, 0) the kinetic operator
 * H = -c1 D^2/2M is solved by the selected solver; any non-zero c2..c4 adds
 * the reference's magnetic, E-derivative and spin-orbit terms
 * (nrqcd.f90:1244-1381), solved by FP64 GMRES on spin-coupled columns.
 * stats may be NULL. */
int nrq_heavy_update(nrq_heavy *h, const nrq_gauge *g, double mass, const double c[8],
                     double stepwidth, nrq_solve_stats *stats);
/* nrq_heavy_update for n heavy fields on one gauge configuration, heavy field
 * k with mass[k] (same Wilson coefficients and step width): all their live
 * colour fields (at most 48) are right-hand sides of one batched solve that
 * streams the links once per chunk of fields.  Results equal n separate
 * nrq_heavy_update calls to the solver tolerance.  No reference counterpart:
 * the reference updates one NRQCDField per call (nrqcd.f90:1053). */
int nrq_heavy_update_batch(nrq_heavy *const *h, const double *mass, int n, const nrq_gauge *g, const double c[8],
                           double stepwidth, nrq_solve_stats *stats);
/* GetNorm_Quark / GetNorm_AntiQ, nrqcd.f90:547-611 */
int nrq_heavy_norms(const nrq_heavy *h, double *norm_quark, double *norm_antiq);
/* GetMesoncorrelator_{O1_1S0, O1_3S1, O8_1S0, O8_3S1}_ZeroMomentum,
 * nrqcd.f90:1484-1661; out = 4 complex numbers in that order */
int nrq_heavy_correlators(const nrq_heavy *h, double out[8]);
/* GetMesoncorrelator_{O1_1S0, O1_3S1, O8_1S0, O8_3S1}_ZeroMomentum_gconjg,
 * nrqcd.f90:1728-1905 (conj(A) in place of A^dag); same order as above */
int nrq_heavy_correlators_gconjg(const nrq_heavy *h, double out[8]);
/* nrq_heavy_correlators without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device or pinned memory
 * (graph-capturing or tight serving loops, harnesses that forbid hipMalloc):
 * the plain correlator of nrq_heavy_correlators computed by the same kernels
 * with the partial sums in a fixed 524,416-byte device array of the library
 * (one per GPU, held until the process ends) instead of the solver workspace
 * that nrq_heavy_correlators reserves on first use.  Any lattice size fits
 * (the trace kernel strides over the sites); a result is returned to the
 * host, so the call synchronises the library stream.  Calls that overlap on
 * one GPU share that array and race. */
int nrq_heavy_correlators_noalloc(const nrq_heavy *h, double out[8]);

/* ---- spectral functions (main.f90 ComputeSpectrum, twfft.f90) --------- */

/* ComputeSpectrum, main.f90:1865-2100 (Symmetrisation :1970-1992,
 * ApplyHannWindow windowing.f90:11-25, Symm2FFTformat :2000-2016, t2w
 * twfft.f90:73-105, spectrum = 2 Im), for `batch` correlators at once.
 * signals: batch x n complex (t = 0, dt, ..., (n-1) dt; n >= 2);
 * spectra: batch x (2n - 2) reals in transform order (index 0: omega = 0,
 * index l: omega = 2 pi l / (dt (2n - 2)), upper half negative frequencies);
 * stages: NULL, or batch x 3 x (2n - 2) complex receiving the symmetrised,
 * the Hann-windowed and the reordered signal of every correlator (the
 * reference's symm.txt, hanned.txt and fftformat.txt). */
int nrq_spectral_functions(const double *signals, int64_t n, int64_t batch, double dt, double *spectra,
                           double *stages);
/* t2w, twfft.f90:73-105: in place, F_l = dt sum_j f_j exp(+2 pi i j l / n) */
int nrq_t2w(double *data, int64_t n, double dt);
/* w2t, twfft.f90:33-66: in place, f_j = sum_l F_l exp(-2 pi i j l / n) / (dt n) */
int nrq_w2t(double *data, int64_t n, double dt);

/* ---- learned preconditioner ------------------------------------------
 * No reference counterpart (the reference uses PETSc's ILU(0)/block-Jacobi,
 * nrqcd.f90:786-804).  A 4-layer 3-D convolutional network N, 32 channels
 * wide, 3^3 periodic kernels, ReLU, no biases, BF16 weights and activations
 * with FP32 accumulation on matrix cores.  Its 18 real input channels per
 * site are one colour-vector field v (3 colours x re/im) and its covariant
 * features K v / kmax and K^2 v / kmax^2, with K the covariant Laplacian of
 * the CN operator 1 + beta K: the inputs transform like v under gauge
 * transformations.  FGMRES uses z = v - beta K v + (beta K)^2 v
 * - (beta kmax)^3 N(v, K v / kmax, K^2 v / kmax^2) (output channels 0..5,
 * kmax = sum_i 2 / a_i^2) as its right preconditioner: the second-order
 * Neumann polynomial of (1 + beta K)^-1 corrected by the network in units of
 * the polynomial's leading remainder, so that one network serves every mass
 * (with all weights 0, the polynomial alone).  The lattice x extent must be
 * a multiple of 16. */
typedef struct nrq_cnn nrq_cnn;
#define NRQ_CNN_NWEIGHTS (4 * 27 * 32 * 32) /* [layer][tap][co][ci], tap = 9(dz+1)+3(dy+1)+(dx+1);
                                               channels past 18 of layer 1 inputs and
                                               layer 4 outputs are zero */
/* He-initialised weights (the last layer scaled by 0.1) from `seed` */
int nrq_cnn_create(nrq_cnn **net, uint64_t seed);
int nrq_cnn_destroy(nrq_cnn *net);
int nrq_cnn_get_weights(const nrq_cnn *net, float *w);
int nrq_cnn_set_weights(nrq_cnn *net, const float *w);
/* out = N(in) on the current lattice; in, out: [site][18] floats (host) */
int nrq_cnn_forward(nrq_cnn *net, const float *in, float *out);
/* Mean device time of one network apply (four convolution layers, one colour
 * field with its features, inputs resident on the device) over `reps`
 * applies after one warm-up, by HIP events on the library stream.  For
 * reporting the network against the matrix-core peak; no reference
 * counterpart. */
int nrq_cnn_time_forward(nrq_cnn *net, int reps, double *ms);
/* loss = sum (N(in) - target)^2 / norm and its gradient with respect to the
 * NRQ_CNN_NWEIGHTS weights (host arrays, [site][18] layouts) */
int nrq_cnn_loss_grad(nrq_cnn *net, const float *in, const float *target, double norm, double *loss, float *grad);
/* nrq_cnn_loss_grad that also returns the network's own BF16 values (as
 * floats, [layer][site][32]): acts[l] the input of layer l, deltas[l] the
 * loss gradient at the output of layer l, from which that layer's weight
 * gradient is formed.  Either may be NULL.  For checking the kernels. */
int nrq_cnn_loss_grad_trace(nrq_cnn *net, const float *in, const float *target, double norm, double *loss, float *grad,
                            float *acts, float *deltas);
/* Adds training pairs from `nsolves` plain FP64 GMRES solves (true relative
 * residual 1e-12) of CN systems on g at `mass` and `stepwidth` (c1 = 1):
 * each solve takes 12 random Gaussian fields psi (6 quark, 6 antiquark),
 * forms y = S psi and solves S^+ x = y; every field gives one pair: input
 * [y, K y / kmax, K^2 y / kmax^2] / |y| (kmax = sum_i 2 / a_i^2), target
 * (x - (y - beta K y + (beta K)^2 y)) / (-(beta kmax)^3 |y|) (what the
 * Neumann polynomial misses of the zero guess's error, in units of its
 * leading term) in channels 0..5 and zero in 6..17: the same functions of y
 * at every mass.  All pairs of one network come from one lattice size. */
int nrq_cnn_harvest(nrq_cnn *net, const nrq_gauge *g, double mass, double stepwidth, int nsolves, uint64_t seed);
int nrq_cnn_dataset_size(const nrq_cnn *net, int64_t *count);
/* `epochs` passes of Adam (learning rate lr, one pair per step, order
 * shuffled from seed); mean_loss (may be NULL) = mean relative loss of the
 * last pass */
int nrq_cnn_train(nrq_cnn *net, int epochs, double lr, uint64_t seed, double *mean_loss);
/* network used by NRQ_SOLVER_FGMRES_CNN (NULL: none) */
int nrq_cn_set_preconditioner(nrq_cnn *net);

#ifdef __cplusplus
}
#endif
#endif
