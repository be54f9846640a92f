/*
 * nrqcd_oracle.h -- CPU FP64 restatement of the reference's real-time SU(3)
 * leapfrog and NRQCD Crank-Nicolson path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (nrqcd_amd/) never links, loads or calls it.
 *
 * Storage follows the reference's Fortran (column-major) arrays on one rank,
 * where memory index == lattice index:
 *   links(3,3,3,V) complex128 : U_i(x)_{rc} at r + 3c + 9(i-1) + 27(x-1)
 *                                (gaugeconfiguration_su3.f90:35,134-137)
 *   efield(8,3,V)  real64     : E^a_i(x) at (a-1) + 8(i-1) + 24(x-1)
 *   prop(6,6,V)    complex128 : P_{rc}(x) at r + 6c + 36(x-1)   (nrqcd.f90:60-63)
 * Lattice indices are 1-based, x fastest (lattice.f90:883-931).
 *
 * Pinning status of each function is stated in the header comment of
 * nrqcd_oracle.c.
 */
#ifndef NRQCD_ORACLE_H
#define NRQCD_ORACLE_H
#include <stdint.h>
#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex zc;

/* lattice.f90 */
int     orc_lattice_init(const int64_t ext[3], const double spacing[4]);
int64_t orc_lattice_size(void);
int64_t orc_index(const int64_t pos[3]);
void    orc_position(int64_t idx, int64_t pos[3]);
int64_t orc_neib(int dir, int64_t idx);
int64_t orc_negative_index(int64_t idx);
void    orc_momentum(int64_t idx, zc p[3]);
double  orc_norm2_momentum(int64_t idx);
void    orc_polarisation(const zc p[3], zc eps[9]);

/* su3.f90 / matrixoperations.f90 */
void    orc_generators(zc T[72]);
void    orc_algebra_matrix(const double alpha[8], zc M[9]);
double  orc_algebra_coordinate(int a, const zc M[9]);
zc      orc_trace_with_generator(int a, const zc M[9]);
void    orc_eigh3(const zc H[9], double lam[3], zc V[9]);
void    orc_group_exp(const double alpha[8], zc U[9]);
void    orc_pauli(int k, zc s[4]);
void    orc_kron(const zc *A, int ra, int ca, const zc *B, int rb, int cb, zc *AB);

/* ranlxd_generator.f90 / random.f90 */
int     orc_ranlxd_init(int level, int64_t seed);
void    orc_ranlxd(double *r, int64_t n);
void    orc_ranlxd_get(int64_t state[25]);
void    orc_normal_cmplx(zc *out, int64_t n);

/* xpfft.f90 conventions: forward e^{-i p x} scaled by a1 a2 a3, backward
 * e^{+i p x} scaled by 1/(N1 a1 N2 a2 N3 a3) */
void    orc_x2p(zc *data);
void    orc_p2x(zc *data);

/* gaugeconfiguration_su3.f90 */
void    orc_cold_init(zc *links, double *efield);
int     orc_box_init(double qs, double amplitude, double coupling, zc *links, double *efield);
void    orc_update_links(zc *links, const double *efield, double stepwidth);
void    orc_update_efield(const zc *links, double *efield, double stepwidth);
void    orc_leapfrog(zc *links, double *efield, double stepwidth, int64_t nsteps);
double  orc_energy(const zc *links, const double *efield);
double  orc_gauss_ref(const zc *links, const double *efield);
double  orc_gauss_max(const zc *links, const double *efield);
void    orc_plaquette(const zc *links, int i, int j, int64_t idx, zc P[9]);
double  orc_mean_plaquette(const zc *links);

/* nrqcd.f90 */
void    orc_heavy_init_point(zc *prop, int64_t idx);
void    orc_apply_system(const zc *links, const zc *in, zc *out, double mass,
                         zc c1, double stepwidth, int hermitian_conjugate);
typedef struct {
    int64_t iterations;      /* total Krylov iterations over all columns */
    int64_t solves;          /* columns with non-zero right-hand side */
    double  max_true_relres; /* worst FP64 true relative residual */
} orc_solve_stats;
int     orc_cn_update(const zc *links, zc *quark, zc *antiq, double mass,
                      const zc c[4], double stepwidth, double tol, int restart,
                      int64_t maxit, orc_solve_stats *stats);
double  orc_prop_norm(const zc *prop);
zc      orc_corr_o1_1s0(const zc *quark, const zc *antiq);
zc      orc_corr_o1_3s1(const zc *quark, const zc *antiq);
zc      orc_corr_o8_1s0(const zc *quark, const zc *antiq);
zc      orc_corr_o8_3s1(const zc *quark, const zc *antiq);
zc      orc_corr_o1_1s0_gconjg(const zc *quark, const zc *antiq);
zc      orc_corr_o1_3s1_gconjg(const zc *quark, const zc *antiq);
zc      orc_corr_o8_1s0_gconjg(const zc *quark, const zc *antiq);
zc      orc_corr_o8_3s1_gconjg(const zc *quark, const zc *antiq);

int     orc_num_threads(void);

/* Planted mistakes, for the tests that show a check fails when the oracle
 * is wrong (0, the default, plants nothing; every other use leaves it 0):
 *   1  staple coefficient (a0/a_i)^2 instead of (a0/a_k)^2
 *   2  magnetic energy divided by a_i a_i instead of a_i a_j
 *   3  box IC: A amplitude sqrt(2) too large
 *   4  box IC: |p| powers of the A and E amplitudes swapped
 *   5  box IC: E takes the noise of the two polarisations swapped
 *   6  box IC: E amplitude real (the factor i dropped) */
void    orc_set_plant(int which);

#ifdef __cplusplus
}
#endif
#endif
