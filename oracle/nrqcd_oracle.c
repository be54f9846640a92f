/*
 * nrqcd_oracle.c -- CPU FP64 restatement of the reference's real-time SU(3)
 * classical-statistical leapfrog and NRQCD Crank-Nicolson step.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product library never does.
 *
 * Every function cites the reference file:line it restates.  How each one is
 * pinned (checked by something other than this file):
 *   lattice index/neighbour maps  closed form + brute-force enumeration on odd
 *                                 and even lattices (tests/test_oracle_lattice.py)
 *   generators, algebra maps, exp pinned against the reference's own su3 /
 *                                 matrixoperations modules built in oracle/_ref
 *                                 and against scipy.linalg.expm
 *   ranlxd, Box-Muller            pinned bit-exactly against the reference's
 *                                 ranlxd_generator built in oracle/_ref; the
 *                                 Box-Muller formula against random.f90:238-253
 *                                 evaluated in numpy on the _ref stream
 *   x2p / p2x                     numpy.fft with the stated scales
 *   box initial condition         numpy restatement on 4^3 using the _ref
 *                                 stream; spectrum statistics
 *   leapfrog, energy, Gauss       invariants: time reversibility, gauge
 *                                 covariance, unitarity, energy drift O(dt^2),
 *                                 Gauss constraint conservation, single-link
 *                                 hand-computable case
 *   CN step, norms, correlators   free-field closed form in momentum space,
 *                                 Hermiticity by dense comparison, gauge
 *                                 covariance, norm conservation, reversibility
 *   polarisation vectors          cross-product form (eps1 = z x p / |p12|
 *                                 or -x x p / |p23|, eps2 = conj(eps1) x p^),
 *                                 transversality, norms, orthogonality
 */
#include "nrqcd_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double PI = 3.14159265358979323846264338327950288;

/* ------------------------------------------------------------------ */
/* lattice (lattice.f90)                                              */
/* ------------------------------------------------------------------ */
static int64_t N[3] = {0, 0, 0};
static int64_t VOL = 0;
static double SP[4] = {1, 1, 1, 1};
static const double ZERO_TOL = 1e-12; /* tolerances.f90:23,26 */

/* planted mistake (see nrqcd_oracle.h); 0 plants nothing */
static int PLANT = 0;
void orc_set_plant(int which) { PLANT = which; }

int orc_lattice_init(const int64_t ext[3], const double spacing[4])
{
    for (int i = 0; i < 3; ++i) {
        if (ext[i] < 1) return -1;
        N[i] = ext[i];
    }
    for (int i = 0; i < 4; ++i) SP[i] = spacing[i];
    VOL = N[0] * N[1] * N[2];
    return 0;
}

int64_t orc_lattice_size(void) { return VOL; }

/* GetIndex_FromPosition, lattice.f90:883-898 */
int64_t orc_index(const int64_t pos[3])
{
    return 1 + (pos[0] - 1) + N[0] * (pos[1] - 1) + N[0] * N[1] * (pos[2] - 1);
}

/* GetPosition_FromIndex, lattice.f90:912-931 */
void orc_position(int64_t idx, int64_t pos[3])
{
    int64_t r = idx - 1;
    pos[0] = r % N[0] + 1;
    r /= N[0];
    pos[1] = r % N[1] + 1;
    pos[2] = r / N[1] + 1;
}

/* GetPeriodicPosition, lattice.f90:939-947 (Fortran modulo: result >= 0) */
static int64_t periodic(int64_t p, int64_t n)
{
    int64_t m = (p - 1) % n;
    if (m < 0) m += n;
    return m + 1;
}

/* GetNeib_FromIndex, lattice.f90:846-874 */
int64_t orc_neib(int dir, int64_t idx)
{
    if (dir == 0) return idx;
    int64_t pos[3];
    orc_position(idx, pos);
    int d = abs(dir) - 1;
    pos[d] = periodic(pos[d] + (dir > 0 ? 1 : -1), N[d]);
    return orc_index(pos);
}

/* GetNegativeLatticeIndex / GetNegativeLatticePosition, lattice.f90:495-525 */
int64_t orc_negative_index(int64_t idx)
{
    int64_t pos[3];
    orc_position(idx, pos);
    for (int d = 0; d < 3; ++d) {
        int64_t m = (N[d] - (pos[d] - 1)) % N[d];
        if (m < 0) m += N[d];
        pos[d] = 1 + m;
    }
    return orc_index(pos);
}

/* GetMomentum_BackwardDerivative, lattice.f90:963-979 */
void orc_momentum(int64_t idx, zc p[3])
{
    int64_t pos[3];
    orc_position(idx, pos);
    for (int d = 0; d < 3; ++d) {
        double arg = PI * (double)(pos[d] - 1) / (double)N[d];
        double mag = 2.0 * sin(arg) / SP[d + 1];
        p[d] = mag * (cos(arg) - I * sin(arg));
    }
}

/* GetNorm2Momentum_G = norm2(abs(p)), lattice.f90:1048-1058 */
double orc_norm2_momentum(int64_t idx)
{
    zc p[3];
    orc_momentum(idx, p);
    double s = 0;
    for (int d = 0; d < 3; ++d) s += cabs(p[d]) * cabs(p[d]);
    return sqrt(s);
}

/* GetPolarisationVectors, lattice.f90:1120-1176.  eps[i + 3*pol]. */
void orc_polarisation(const zc p[3], zc eps[9])
{
    double ap = sqrt(cabs(p[0]) * cabs(p[0]) + cabs(p[1]) * cabs(p[1]) + cabs(p[2]) * cabs(p[2]));
    if (ap < ZERO_TOL) {
        for (int k = 0; k < 9; ++k) eps[k] = 0;
        return;
    }
    if (cabs(p[0]) > ZERO_TOL) {
        double n12 = sqrt(cabs(p[0]) * cabs(p[0]) + cabs(p[1]) * cabs(p[1]));
        eps[0] = -p[1] / n12;
        eps[1] = +p[0] / n12;
        eps[2] = 0;
        eps[3] = conj(p[0]) * p[2] / (ap * n12);
        eps[4] = conj(p[1]) * p[2] / (ap * n12);
        eps[5] = -n12 / ap;
    } else {
        double n23 = sqrt(cabs(p[1]) * cabs(p[1]) + cabs(p[2]) * cabs(p[2]));
        eps[0] = 0;
        eps[1] = +p[2] / n23;
        eps[2] = -p[1] / n23;
        eps[3] = 1;
        eps[4] = 0;
        eps[5] = 0;
    }
    for (int d = 0; d < 3; ++d) eps[6 + d] = p[d] / ap;
}

/* ------------------------------------------------------------------ */
/* SU(3) algebra (su3.f90, matrixoperations.f90)                      */
/* ------------------------------------------------------------------ */
/* 3x3 matrices are column-major: M[r + 3c]. */
static zc GEN[8][9];
static int gen_ready = 0;

/* Generators T^a = lambda^a/2 in the reference's convention, su3.f90:97-131.
 * The reshape there fills column-major, which makes T^2, T^5, T^7 the complex
 * conjugates of the textbook Gell-Mann matrices. */
static void init_generators(void)
{
    if (gen_ready) return;
    memset(GEN, 0, sizeof(GEN));
#define SETG(a, r, c, v) GEN[(a) - 1][((r) - 1) + 3 * ((c) - 1)] = (v)
    SETG(1, 2, 1, 0.5); SETG(1, 1, 2, 0.5);
    SETG(2, 2, 1, -0.5 * I); SETG(2, 1, 2, +0.5 * I);
    SETG(3, 1, 1, 0.5); SETG(3, 2, 2, -0.5);
    SETG(4, 3, 1, 0.5); SETG(4, 1, 3, 0.5);
    SETG(5, 3, 1, -0.5 * I); SETG(5, 1, 3, +0.5 * I);
    SETG(6, 3, 2, 0.5); SETG(6, 2, 3, 0.5);
    SETG(7, 3, 2, -0.5 * I); SETG(7, 2, 3, +0.5 * I);
    SETG(8, 1, 1, 0.5 / sqrt(3.0)); SETG(8, 2, 2, 0.5 / sqrt(3.0)); SETG(8, 3, 3, -1.0 / sqrt(3.0));
#undef SETG
    gen_ready = 1;
}

void orc_generators(zc T[72])
{
    init_generators();
    memcpy(T, GEN, sizeof(GEN));
}

static void mm3(const zc *A, const zc *B, zc *C)
{
    zc t[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            zc s = 0;
            for (int k = 0; k < 3; ++k) s += A[r + 3 * k] * B[k + 3 * c];
            t[r + 3 * c] = s;
        }
    memcpy(C, t, sizeof(t));
}

static void dag3(const zc *A, zc *B)
{
    zc t[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) t[r + 3 * c] = conj(A[c + 3 * r]);
    memcpy(B, t, sizeof(t));
}

/* GetAlgebraMatrix, su3.f90:287-301 */
void orc_algebra_matrix(const double alpha[8], zc M[9])
{
    init_generators();
    for (int k = 0; k < 9; ++k) M[k] = 0;
    for (int a = 0; a < 8; ++a)
        for (int k = 0; k < 9; ++k) M[k] += alpha[a] * GEN[a][k];
}

/* GetTraceWithGenerator, su3.f90:369-381: Tr(T^a M) */
zc orc_trace_with_generator(int a, const zc M[9])
{
    init_generators();
    zc s = 0;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) s += GEN[a - 1][r + 3 * c] * M[c + 3 * r];
    return s;
}

/* GetAlgebraCoordinate, su3.f90:349-357: 2 Re Tr(T^a M) */
double orc_algebra_coordinate(int a, const zc M[9])
{
    return 2.0 * creal(orc_trace_with_generator(a, M));
}

/* Hermitian eigen-decomposition H = V diag(lam) V^dagger by cyclic complex
 * Jacobi rotations; the same method as EigenH, matrixoperations.f90:187-312. */
void orc_eigh3(const zc H[9], double lam[3], zc V[9])
{
    zc A[9];
    memcpy(A, H, sizeof(A));
    for (int k = 0; k < 9; ++k) V[k] = 0;
    V[0] = V[4] = V[8] = 1;
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = 0, dia = 0;
        for (int q = 1; q < 3; ++q)
            for (int p = 0; p < q; ++p) off += creal(A[p + 3 * q] * conj(A[p + 3 * q]));
        for (int p = 0; p < 3; ++p) dia += creal(A[4 * p]) * creal(A[4 * p]);
        if (off <= 1e-64 * dia || off == 0.0 || off < 1e-300) break;
        for (int q = 1; q < 3; ++q)
            for (int p = 0; p < q; ++p) {
                zc apq = A[p + 3 * q];
                double m = cabs(apq);
                if (m == 0.0) continue;
                zc e = apq / m;
                double app = creal(A[4 * p]), aqq = creal(A[4 * q]);
                double th = (aqq - app) / (2.0 * m);
                double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                /* J = [[c, s], [-s conj(e), c conj(e)]] on coordinates (p,q) */
                zc jpp = c, jpq = s, jqp = -s * conj(e), jqq = c * conj(e);
                for (int k = 0; k < 3; ++k) { /* A <- A J */
                    zc akp = A[k + 3 * p], akq = A[k + 3 * q];
                    A[k + 3 * p] = akp * jpp + akq * jqp;
                    A[k + 3 * q] = akp * jpq + akq * jqq;
                }
                for (int k = 0; k < 3; ++k) { /* A <- J^dagger A */
                    zc apk = A[p + 3 * k], aqk = A[q + 3 * k];
                    A[p + 3 * k] = conj(jpp) * apk + conj(jqp) * aqk;
                    A[q + 3 * k] = conj(jpq) * apk + conj(jqq) * aqk;
                }
                A[p + 3 * q] = 0;
                A[q + 3 * p] = 0;
                A[4 * p] = creal(A[4 * p]);
                A[4 * q] = creal(A[4 * q]);
                for (int k = 0; k < 3; ++k) { /* V <- V J */
                    zc vkp = V[k + 3 * p], vkq = V[k + 3 * q];
                    V[k + 3 * p] = vkp * jpp + vkq * jqp;
                    V[k + 3 * q] = vkp * jpq + vkq * jqq;
                }
            }
    }
    for (int p = 0; p < 3; ++p) lam[p] = creal(A[4 * p]);
}

/* GetGroupExp_fromAlgebraCoordinates -> ExpAH_diaglibrary,
 * su3.f90:235-274 and matrixoperations.f90:547-567: exp(i alpha^a T^a). */
void orc_group_exp(const double alpha[8], zc U[9])
{
    zc H[9], V[9];
    double lam[3];
    orc_algebra_matrix(alpha, H);
    orc_eigh3(H, lam, V);
    zc ph[3];
    for (int k = 0; k < 3; ++k) ph[k] = cos(lam[k]) + I * sin(lam[k]);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            zc s = 0;
            for (int k = 0; k < 3; ++k) s += V[r + 3 * k] * ph[k] * conj(V[c + 3 * k]);
            U[r + 3 * c] = s;
        }
}

/* GetKronProd, matrixoperations.f90:81-104 (column-major operands) */
void orc_kron(const zc *A, int ra, int ca, const zc *B, int rb, int cb, zc *AB)
{
    int ld = ra * rb;
    for (int I1 = 0; I1 < ra; ++I1)
        for (int J1 = 0; J1 < ca; ++J1)
            for (int i = 0; i < rb; ++i)
                for (int j = 0; j < cb; ++j)
                    AB[(I1 * rb + i) + ld * (J1 * cb + j)] = A[I1 + ra * J1] * B[i + rb * j];
}

/* ------------------------------------------------------------------ */
/* RANLUX ranlxd (ranlxd_generator.f90, Luscher's ranlxd v2.1)        */
/* ------------------------------------------------------------------ */
static struct {
    int64_t pr, ir, jr, ir_old, init;
    double one_bit, carry;
    double x[12];
} RX;

/* ranlxd_init, ranlxd_generator.f90:251-313 */
int orc_ranlxd_init(int level, int64_t seed)
{
    if (level == 1) RX.pr = 202;
    else if (level == 2) RX.pr = 397;
    else return -1;
    int64_t bits[31], s = seed;
    for (int k = 0; k < 31; ++k) { bits[k] = s % 2; s /= 2; }
    if (seed <= 0 || s != 0) return -2;
    RX.one_bit = ldexp(1.0, -48);
    int ib = 0, jb = 18;
    for (int k = 0; k < 12; ++k) {
        double v = 0;
        for (int l = 0; l < 48; ++l) {
            v = v + v + (double)((bits[ib] + 1) % 2);
            bits[ib] = (bits[ib] + bits[jb]) % 2;
            ib = (ib + 1) % 31;
            jb = (jb + 1) % 31;
        }
        RX.x[k] = RX.one_bit * v;
    }
    RX.carry = 0;
    RX.ir = 11;
    RX.jr = 7;
    RX.ir_old = 0;
    RX.init = 1;
    return 0;
}

/* update, ranlxd_generator.f90:77-224: pr steps of the subtract-with-borrow
 * recurrence x[ir] <- x[jr] - x[ir] - carry (the reference unrolls the middle
 * part by 12; the sequence of operations is the same). */
static void rx_update(void)
{
    for (int64_t k = 0; k < RX.pr; ++k) {
        double y = RX.x[RX.jr] - RX.x[RX.ir] - RX.carry;
        if (y < 0) { RX.carry = RX.one_bit; y += 1.0; }
        else RX.carry = 0;
        RX.x[RX.ir] = y;
        RX.ir = (RX.ir + 1) % 12;
        RX.jr = (RX.jr + 1) % 12;
    }
    RX.ir_old = RX.ir;
}

/* ranlxd, ranlxd_generator.f90:319-349 */
void orc_ranlxd(double *r, int64_t n)
{
    if (!RX.init) orc_ranlxd_init(1, 1);
    for (int64_t k = 0; k < n; ++k) {
        RX.ir = (RX.ir + 1) % 12;
        if (RX.ir == RX.ir_old) rx_update();
        r[k] = RX.x[RX.ir];
    }
}

/* ranlxd_get, ranlxd_generator.f90:355-382 */
void orc_ranlxd_get(int64_t state[25])
{
    double sbase = ldexp(1.0, 24), base = ldexp(1.0, 48);
    for (int k = 0; k < 12; ++k) {
        double x = sbase * RX.x[k];
        double y2 = trunc(x);
        double y1 = sbase * (x - y2);
        state[2 * k] = (int64_t)y1;
        state[2 * k + 1] = (int64_t)y2;
    }
    int64_t k = 12 * RX.pr + RX.ir;
    k = 12 * k + RX.jr;
    k = 12 * k + RX.ir_old;
    state[24] = 2 * k + (int64_t)(RX.carry * base);
}

/* GetRandomNormalCmplxArray, random.f90:238-253 (Box-Muller on pairs) */
void orc_normal_cmplx(zc *out, int64_t n)
{
    double *r = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    orc_ranlxd(r, 2 * n);
    const double twopi = 2.0 * PI;
    for (int64_t i = 0; i < n; ++i) {
        double rad = sqrt(-2.0 * log(r[2 * i]));
        out[i] = rad * (cos(twopi * r[2 * i + 1]) + I * sin(twopi * r[2 * i + 1]));
    }
    free(r);
}

/* ------------------------------------------------------------------ */
/* 3-D FFT with the xpfft.f90:129-137 scale and sign conventions      */
/* ------------------------------------------------------------------ */
static void fft1d(zc *v, int64_t n, int sign, zc *tmp, const zc *w)
{
    /* w[k] = exp(sign*2*pi*i*k/n), k < n */
    if ((n & (n - 1)) == 0) {
        for (int64_t i = 1, j = 0; i < n; ++i) { /* bit reversal */
            int64_t bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) { zc t = v[i]; v[i] = v[j]; v[j] = t; }
        }
        for (int64_t len = 2; len <= n; len <<= 1) {
            int64_t step = n / len;
            for (int64_t i = 0; i < n; i += len)
                for (int64_t k = 0; k < len / 2; ++k) {
                    zc u = v[i + k], t = w[k * step] * v[i + k + len / 2];
                    v[i + k] = u + t;
                    v[i + k + len / 2] = u - t;
                }
        }
    } else {
        for (int64_t k = 0; k < n; ++k) {
            zc s = 0;
            for (int64_t j = 0; j < n; ++j) s += w[(k * j) % n] * v[j];
            tmp[k] = s;
        }
        memcpy(v, tmp, sizeof(zc) * (size_t)n);
    }
    (void)sign;
}

static void fft3d(zc *data, int sign, double scale)
{
    int64_t maxn = N[0];
    if (N[1] > maxn) maxn = N[1];
    if (N[2] > maxn) maxn = N[2];
    zc *line = (zc *)malloc(sizeof(zc) * (size_t)maxn);
    zc *tmp = (zc *)malloc(sizeof(zc) * (size_t)maxn);
    zc *w = (zc *)malloc(sizeof(zc) * (size_t)maxn);
    int64_t stride[3] = {1, N[0], N[0] * N[1]};
    for (int d = 0; d < 3; ++d) {
        int64_t n = N[d];
        if (n == 1) continue;
        for (int64_t k = 0; k < n; ++k) {
            double ang = 2.0 * PI * (double)k / (double)n;
            w[k] = cos(ang) + I * (double)sign * sin(ang);
        }
        for (int64_t base = 0; base < VOL; ++base) {
            /* base runs over sites whose coordinate d is zero */
            if ((base / stride[d]) % n != 0) continue;
            for (int64_t k = 0; k < n; ++k) line[k] = data[base + k * stride[d]];
            fft1d(line, n, sign, tmp, w);
            for (int64_t k = 0; k < n; ++k) data[base + k * stride[d]] = line[k];
        }
    }
    for (int64_t s = 0; s < VOL; ++s) data[s] *= scale;
    free(line);
    free(tmp);
    free(w);
}

/* x2p: DFTI forward (e^{-i}), FORWARD_SCALE = a1 a2 a3 (xpfft.f90:134,607-622) */
void orc_x2p(zc *data) { fft3d(data, -1, SP[1] * SP[2] * SP[3]); }

/* p2x: DFTI backward (e^{+i}), BACKWARD_SCALE = 1/Volume (xpfft.f90:136,629-645) */
void orc_p2x(zc *data)
{
    double vol = (double)N[0] * SP[1] * (double)N[1] * SP[2] * (double)N[2] * SP[3];
    fft3d(data, +1, 1.0 / vol);
}

/* ------------------------------------------------------------------ */
/* gauge configuration (gaugeconfiguration_su3.f90)                   */
/* ------------------------------------------------------------------ */
#define LNK(U, i, x) ((U) + 9 * ((i) - 1) + 27 * ((x) - 1)) /* i = 1..3, x = 1..V */
#define EFD(E, i, x) ((E) + 8 * ((i) - 1) + 24 * ((x) - 1))

/* ColdInit, gaugeconfiguration_su3.f90:363-378 */
void orc_cold_init(zc *links, double *efield)
{
    for (int64_t x = 1; x <= VOL; ++x)
        for (int i = 1; i <= 3; ++i) {
            zc *u = LNK(links, i, x);
            for (int k = 0; k < 9; ++k) u[k] = 0;
            u[0] = u[4] = u[8] = 1;
            double *e = EFD(efield, i, x);
            for (int a = 0; a < 8; ++a) e[a] = 0;
        }
}

/* GetBoxOccupation, gaugeconfiguration_su3.f90:1455-1471 */
static double box_occupation(double p, double qs, double amp, double g)
{
    return p < qs ? amp : g * g / 2.0;
}

/* TransversePolarisedOccupiedInit_Box -> TransversePolarisedOccupiedInit,
 * gaugeconfiguration_su3.f90:906-951 and 1263-1446, on one rank.  Uses the
 * current ranlxd stream (random.f90:126-140 seeds it). */
int orc_box_init(double qs, double amplitude, double coupling, zc *links, double *efield)
{
    const int64_t V = VOL;
    zc *r = (zc *)calloc((size_t)(V * 16), sizeof(zc));       /* r(V, 8, 2) */
    zc *af = (zc *)calloc((size_t)(V * 24), sizeof(zc));      /* afield(V, 8, 3) */
    zc *ef = (zc *)calloc((size_t)(V * 24), sizeof(zc));
    if (!r || !af || !ef) { free(r); free(af); free(ef); return -1; }
    zc nrm[8];
    /* :1334-1340 -- per site, per polarisation, 8 complex normals, real part kept */
    for (int64_t x = 0; x < V; ++x)
        for (int pol = 0; pol < 2; ++pol) {
            orc_normal_cmplx(nrm, 8);
            for (int a = 0; a < 8; ++a) r[x + V * (a + 8 * pol)] = creal(nrm[a]);
        }
    /* :1343-1347 */
    for (int pol = 0; pol < 2; ++pol)
        for (int a = 0; a < 8; ++a) orc_x2p(r + V * (a + 8 * pol));
    /* :1353-1381 */
    for (int64_t x = 0; x < V; ++x) {
        double pn = orc_norm2_momentum(x + 1);
        if (!(pn > ZERO_TOL)) continue;
        zc p[3], eps[9];
        orc_momentum(x + 1, p);
        orc_polarisation(p, eps);
        double f = box_occupation(pn, qs, amplitude, coupling);
        double pa = sqrt(f / pn * 2.0);
        zc pe = I * sqrt(f * pn * 2.0);
        if (PLANT == 3) pa *= sqrt(2.0);
        if (PLANT == 4) {
            pa = sqrt(f * pn * 2.0);
            pe = I * sqrt(f / pn * 2.0);
        }
        if (PLANT == 6) pe = sqrt(f * pn * 2.0);
        for (int a = 0; a < 8; ++a) {
            zc r1 = r[x + V * a], r2 = r[x + V * (a + 8)];
            for (int i = 0; i < 3; ++i) {
                zc mode = r1 * eps[i] + r2 * eps[i + 3];
                af[x + V * (a + 8 * i)] = pa * mode;
                ef[x + V * (a + 8 * i)] = pe * (PLANT == 5 ? r2 * eps[i] + r1 * eps[i + 3] : mode);
            }
        }
    }
    /* :1420-1425 */
    for (int i = 0; i < 3; ++i)
        for (int a = 0; a < 8; ++a) {
            orc_p2x(af + V * (a + 8 * i));
            orc_p2x(ef + V * (a + 8 * i));
        }
    /* :1431-1443 */
    for (int64_t x = 0; x < V; ++x)
        for (int i = 1; i <= 3; ++i) {
            double al[8];
            for (int a = 0; a < 8; ++a) al[a] = creal(af[x + V * (a + 8 * (i - 1))]) * SP[i];
            orc_group_exp(al, LNK(links, i, x + 1));
            double *e = EFD(efield, i, x + 1);
            for (int a = 0; a < 8; ++a) e[a] = creal(ef[x + V * (a + 8 * (i - 1))]) * SP[0] * SP[i];
        }
    free(r);
    free(af);
    free(ef);
    return 0;
}

/* Update_Links_Leapfrog, gaugeconfiguration_su3.f90:2118-2140 */
void orc_update_links(zc *links, const double *efield, double stepwidth)
{
#pragma omp parallel for schedule(static)
    for (int64_t x = 1; x <= VOL; ++x)
        for (int i = 1; i <= 3; ++i) {
            double al[8];
            const double *e = EFD(efield, i, x);
            for (int a = 0; a < 8; ++a) al[a] = e[a] * stepwidth;
            zc T[9];
            orc_group_exp(al, T);
            mm3(T, LNK(links, i, x), LNK(links, i, x));
        }
}

/* GetUStaple, gaugeconfiguration_su3.f90:2208-2228:
 * U_k(x+i) U_i(x+k)^dag U_k(x)^dag */
static void ustaple(const zc *U, int i, int k, int64_t x, zc *S)
{
    int64_t xi = orc_neib(+i, x), xk = orc_neib(+k, x);
    zc a[9], b[9];
    dag3(LNK(U, i, xk), a);
    mm3(LNK(U, k, xi), a, S);
    dag3(LNK(U, k, x), b);
    mm3(S, b, S);
}

/* GetDStaple, gaugeconfiguration_su3.f90:2236-2258:
 * U_k(x+i-k)^dag U_i(x-k)^dag U_k(x-k) */
static void dstaple(const zc *U, int i, int k, int64_t x, zc *S)
{
    int64_t xmk = orc_neib(-k, x), ximk = orc_neib(+i, xmk);
    zc a[9], b[9];
    dag3(LNK(U, k, ximk), a);
    dag3(LNK(U, i, xmk), b);
    mm3(a, b, S);
    mm3(S, LNK(U, k, xmk), S);
}

/* Update_Efield_Leapfrog, gaugeconfiguration_su3.f90:2147-2199 */
void orc_update_efield(const zc *links, double *efield, double stepwidth)
{
    init_generators();
#pragma omp parallel for schedule(static)
    for (int64_t x = 1; x <= VOL; ++x)
        for (int i = 1; i <= 3; ++i) {
            zc sum[9] = {0};
            for (int k = 1; k <= 3; ++k) {
                if (k == i) continue;
                zc su[9], sd[9];
                ustaple(links, i, k, x, su);
                dstaple(links, i, k, x, sd);
                const int kc = PLANT == 1 ? i : k;
                double coef = stepwidth * (SP[0] / SP[kc]) * (SP[0] / SP[kc]);
                for (int m = 0; m < 9; ++m) sum[m] += coef * (su[m] + sd[m]);
            }
            zc w[9];
            mm3(LNK(links, i, x), sum, w);
            double *e = EFD(efield, i, x);
            for (int a = 1; a <= 8; ++a) e[a - 1] -= 2.0 * cimag(orc_trace_with_generator(a, w));
        }
}

/* Update_Leapfrog, gaugeconfiguration_su3.f90:2082-2111 (one rank: halo
 * exchanges are no-ops) */
void orc_leapfrog(zc *links, double *efield, double stepwidth, int64_t nsteps)
{
    for (int64_t s = 0; s < nsteps; ++s) {
        if (stepwidth > 0) {
            orc_update_links(links, efield, stepwidth);
            orc_update_efield(links, efield, stepwidth);
        } else {
            orc_update_efield(links, efield, stepwidth);
            orc_update_links(links, efield, stepwidth);
        }
    }
}

/* GetSpatialPlaquette_M, gaugeconfiguration_su3.f90:1591-1615:
 * U_i(x) U_j(x+i) U_i(x+j)^dag U_j(x)^dag */
void orc_plaquette(const zc *links, int i, int j, int64_t x, zc P[9])
{
    zc a[9], b[9];
    mm3(LNK(links, i, x), LNK(links, j, orc_neib(+i, x)), P);
    dag3(LNK(links, i, orc_neib(+j, x)), a);
    mm3(P, a, P);
    dag3(LNK(links, j, x), b);
    mm3(P, b, P);
}

static double sum_sequential(const double *v, int64_t n)
{
    double s = 0;
    for (int64_t k = 0; k < n; ++k) s += v[k];
    return s;
}

/* GetEnergy, gaugeconfiguration_su3.f90:1769-1817 */
double orc_energy(const zc *links, const double *efield)
{
    double *site = (double *)malloc(sizeof(double) * (size_t)VOL);
#pragma omp parallel for schedule(static)
    for (int64_t x = 1; x <= VOL; ++x) {
        double s = 0;
        for (int i = 1; i <= 3; ++i) {
            const double *e = EFD(efield, i, x);
            double ee = 0;
            for (int a = 0; a < 8; ++a) {
                double v = e[a] / SP[i] / SP[0];
                ee += v * v;
            }
            s += ee / 2.0;
            for (int j = i + 1; j <= 3; ++j) {
                zc P[9];
                orc_plaquette(links, i, j, x, P);
                double tr = creal(P[0] + P[4] + P[8]);
                s += 2.0 * 3.0 * (1.0 - tr / 3.0) / SP[i] / SP[PLANT == 2 ? i : j];
            }
        }
        site[x - 1] = s;
    }
    double r = sum_sequential(site, VOL);
    free(site);
    return r;
}

/* GetDeviationFromGaussLaw, gaugeconfiguration_su3.f90:1480-1531 (sum of
 * 2|Im Tr(T^a D)| -- zero up to rounding by construction) */
double orc_gauss_ref(const zc *links, const double *efield)
{
    double *site = (double *)malloc(sizeof(double) * (size_t)VOL);
#pragma omp parallel for schedule(static)
    for (int64_t x = 1; x <= VOL; ++x) {
        double s = 0;
        for (int i = 1; i <= 3; ++i) {
            int64_t nb = orc_neib(-i, x);
            zc em[9], en[9], ud[9], t[9];
            orc_algebra_matrix(EFD(efield, i, x), em);
            orc_algebra_matrix(EFD(efield, i, nb), en);
            dag3(LNK(links, i, nb), ud);
            mm3(ud, en, t);
            mm3(t, LNK(links, i, nb), t);
            zc d[9];
            for (int m = 0; m < 9; ++m) d[m] = (em[m] - t[m]) / (SP[i] * SP[i]) / SP[0];
            for (int a = 1; a <= 8; ++a) s += 2.0 * fabs(cimag(orc_trace_with_generator(a, d)));
        }
        site[x - 1] = s;
    }
    double r = sum_sequential(site, VOL);
    free(site);
    return r;
}

/* GetG + GetGdeviation, gaugeconfiguration_su3.f90:724-813: max over sites
 * and generators of 2|Tr(T^a G(x))| with G = sum_i (E_i(x) - U_i(x-i)^dag
 * E_i(x-i) U_i(x-i))/a_i/a_0 */
double orc_gauss_max(const zc *links, const double *efield)
{
    double *site = (double *)malloc(sizeof(double) * (size_t)VOL);
#pragma omp parallel for schedule(static)
    for (int64_t x = 1; x <= VOL; ++x) {
        zc G[9] = {0};
        for (int i = 1; i <= 3; ++i) {
            int64_t nb = orc_neib(-i, x);
            zc em[9], en[9], ud[9], t[9];
            orc_algebra_matrix(EFD(efield, i, nb), en);
            orc_algebra_matrix(EFD(efield, i, x), em);
            dag3(LNK(links, i, nb), ud);
            mm3(ud, en, t);
            mm3(t, LNK(links, i, nb), t);
            for (int m = 0; m < 9; ++m) G[m] += (em[m] - t[m]) / SP[i];
        }
        for (int m = 0; m < 9; ++m) G[m] /= SP[0];
        double mx = 0;
        for (int a = 1; a <= 8; ++a) {
            double v = 2.0 * cabs(orc_trace_with_generator(a, G));
            if (v > mx) mx = v;
        }
        site[x - 1] = mx;
    }
    double r = 0;
    for (int64_t k = 0; k < VOL; ++k)
        if (site[k] > r) r = site[k];
    free(site);
    return r;
}

/* Average spatial plaquette (1/3V) sum_{x,i<j} Re Tr P_ij(x) / 3, built from
 * GetPlaquette_M (gaugeconfiguration_su3.f90:1722-1737). */
double orc_mean_plaquette(const zc *links)
{
    double *site = (double *)malloc(sizeof(double) * (size_t)VOL);
#pragma omp parallel for schedule(static)
    for (int64_t x = 1; x <= VOL; ++x) {
        double s = 0;
        for (int i = 1; i <= 3; ++i)
            for (int j = i + 1; j <= 3; ++j) {
                zc P[9];
                orc_plaquette(links, i, j, x, P);
                s += creal(P[0] + P[4] + P[8]) / 3.0;
            }
        site[x - 1] = s;
    }
    double r = sum_sequential(site, VOL) / (3.0 * (double)VOL);
    free(site);
    return r;
}

/* ------------------------------------------------------------------ */
/* NRQCD (nrqcd.f90)                                                  */
/* ------------------------------------------------------------------ */
#define PRP(P, x) ((P) + 36 * ((x) - 1))

/* InitSinglePoint, nrqcd.f90:324-350 (one propagator) */
void orc_heavy_init_point(zc *prop, int64_t idx)
{
    memset(prop, 0, sizeof(zc) * 36 * (size_t)VOL);
    zc *p = PRP(prop, idx);
    for (int d = 0; d < 6; ++d) p[d + 6 * d] = 1.0 / sqrt(6.0);
}

/* BuildSystemMatrix restricted to the c1 term, nrqcd.f90:1175-1242, applied
 * matrix-free to one 6-component column v[(x-1)*6 + d] (d = 2(colour-1) +
 * spin-1, the C2CS = U (x) 1_2 ordering of nrqcd.f90:452-467):
 *   S psi = psi + alpha sum_i [psi(x) - (U_i(x) psi(x+i) + U_i(x-i)^dag psi(x-i))/2]/a_i^2
 * with alpha = -i (stepwidth a_0 / 2) c1 / M.  hermitian_conjugate != 0 applies
 * S^dagger (alpha -> conj(alpha)), the MatHermitianTranspose of nrqcd.f90:1097. */
void orc_apply_system(const zc *links, const zc *in, zc *out, double mass, zc c1,
                      double stepwidth, int hermitian_conjugate)
{
    zc alpha = (-I * (stepwidth / 2.0 * SP[0])) * c1 / mass;
    if (hermitian_conjugate) alpha = conj(alpha);
#pragma omp parallel for schedule(static)
    for (int64_t x = 1; x <= VOL; ++x) {
        zc acc[6] = {0};
        const zc *px = in + 6 * (x - 1);
        for (int i = 1; i <= 3; ++i) {
            double ia2 = 1.0 / (SP[i] * SP[i]);
            int64_t xp = orc_neib(+i, x), xm = orc_neib(-i, x);
            const zc *pp = in + 6 * (xp - 1), *pm = in + 6 * (xm - 1);
            const zc *up = LNK(links, i, x), *um = LNK(links, i, xm);
            for (int c = 0; c < 3; ++c)
                for (int s = 0; s < 2; ++s) {
                    zc hop = 0;
                    for (int k = 0; k < 3; ++k)
                        hop += up[c + 3 * k] * pp[2 * k + s] + conj(um[k + 3 * c]) * pm[2 * k + s];
                    acc[2 * c + s] += (px[2 * c + s] - 0.5 * hop) * ia2;
                }
        }
        for (int d = 0; d < 6; ++d) out[6 * (x - 1) + d] = px[d] + alpha * acc[d];
    }
}

static zc zdot(const zc *a, const zc *b, int64_t n) /* a^dagger b */
{
    zc s = 0;
    for (int64_t k = 0; k < n; ++k) s += conj(a[k]) * b[k];
    return s;
}

static double znrm(const zc *a, int64_t n)
{
    double s = 0;
    for (int64_t k = 0; k < n; ++k) s += creal(a[k]) * creal(a[k]) + cimag(a[k]) * cimag(a[k]);
    return sqrt(s);
}

/* Restarted GMRES(m) with classical Gram-Schmidt and a zero initial guess --
 * the KSP defaults of nrqcd.f90:803-804 (manuals/petsc-3.10.pdf p.87-90)
 * without a preconditioner.  Convergence is judged on the FP64 true residual
 * ||b - A x|| <= tol ||b||. */
typedef struct {
    const zc *links;
    double mass;
    zc c1;
    double stepwidth;
} sysop;

static void opA(const sysop *op, const zc *in, zc *out)
{
    orc_apply_system(op->links, in, out, op->mass, op->c1, op->stepwidth, 1);
}

static int gmres(const sysop *op, const zc *b, zc *x, int64_t n, double tol, int m,
                 int64_t maxit, int64_t *its, double *relres)
{
    double bn = znrm(b, n);
    memset(x, 0, sizeof(zc) * (size_t)n);
    *its = 0;
    if (bn == 0.0) { *relres = 0; return 0; }
    zc *Vb = (zc *)malloc(sizeof(zc) * (size_t)n * (size_t)(m + 1));
    zc *w = (zc *)malloc(sizeof(zc) * (size_t)n);
    zc *H = (zc *)calloc((size_t)((m + 1) * m), sizeof(zc));
    zc *g = (zc *)malloc(sizeof(zc) * (size_t)(m + 1));
    double *cs = (double *)malloc(sizeof(double) * (size_t)m);
    zc *sn = (zc *)malloc(sizeof(zc) * (size_t)m);
    zc *y = (zc *)malloc(sizeof(zc) * (size_t)m);
    int status = 1;
    /* r0 = b (x0 = 0) */
    memcpy(w, b, sizeof(zc) * (size_t)n);
    double rn = bn;
    while (*its < maxit) {
        for (int64_t k = 0; k < n; ++k) Vb[k] = w[k] / rn;
        for (int k = 0; k <= m; ++k) g[k] = 0;
        g[0] = rn;
        int j;
        for (j = 0; j < m && *its < maxit; ++j) {
            zc *vj = Vb + n * j, *vn = Vb + n * (j + 1);
            opA(op, vj, vn);
            ++*its;
            /* classical Gram-Schmidt: all projections from the same vector */
            for (int i = 0; i <= j; ++i) H[i + (m + 1) * j] = zdot(Vb + n * i, vn, n);
            for (int i = 0; i <= j; ++i) {
                zc h = H[i + (m + 1) * j];
                const zc *vi = Vb + n * i;
                for (int64_t k = 0; k < n; ++k) vn[k] -= h * vi[k];
            }
            double hn = znrm(vn, n);
            H[(j + 1) + (m + 1) * j] = hn;
            if (hn > 0)
                for (int64_t k = 0; k < n; ++k) vn[k] /= hn;
            /* apply previous Givens rotations to column j */
            for (int i = 0; i < j; ++i) {
                zc a = H[i + (m + 1) * j], c = H[i + 1 + (m + 1) * j];
                H[i + (m + 1) * j] = cs[i] * a + sn[i] * c;
                H[i + 1 + (m + 1) * j] = -conj(sn[i]) * a + cs[i] * c;
            }
            zc a = H[j + (m + 1) * j];
            double bb = creal(H[j + 1 + (m + 1) * j]);
            double den = sqrt(creal(a * conj(a)) + bb * bb);
            if (den == 0) { cs[j] = 1; sn[j] = 0; }
            else if (cabs(a) == 0) { cs[j] = 0; sn[j] = 1; }
            else {
                cs[j] = cabs(a) / den;
                sn[j] = (a / cabs(a)) * bb / den;
            }
            H[j + (m + 1) * j] = cs[j] * a + sn[j] * bb;
            H[j + 1 + (m + 1) * j] = 0;
            g[j + 1] = -conj(sn[j]) * g[j];
            g[j] = cs[j] * g[j];
            if (cabs(g[j + 1]) <= tol * bn) { ++j; break; }
        }
        /* back substitution H(0:j,0:j) y = g */
        for (int i = j - 1; i >= 0; --i) {
            zc s = g[i];
            for (int k = i + 1; k < j; ++k) s -= H[i + (m + 1) * k] * y[k];
            y[i] = s / H[i + (m + 1) * i];
        }
        for (int i = 0; i < j; ++i)
            for (int64_t k = 0; k < n; ++k) x[k] += y[i] * Vb[n * i + k];
        /* true residual */
        opA(op, x, w);
        for (int64_t k = 0; k < n; ++k) w[k] = b[k] - w[k];
        rn = znrm(w, n);
        *relres = rn / bn;
        if (rn <= tol * bn) { status = 0; break; }
    }
    free(Vb); free(w); free(H); free(g); free(cs); free(sn); free(y);
    return status;
}

/* Update_CrankNicholson, nrqcd.f90:1053-1168, kinetic (c1) Hamiltonian:
 * rhs = S psi, solve S^dagger psi' = rhs, column by column; the antiquark
 * uses -M and conj(c). Returns -3 when c2..c4 are non-zero (not restated). */
int orc_cn_update(const zc *links, zc *quark, zc *antiq, double mass, const zc c[4],
                  double stepwidth, double tol, int restart, int64_t maxit,
                  orc_solve_stats *stats)
{
    if (c[1] != 0 || c[2] != 0 || c[3] != 0) return -3;
    const int64_t n = 6 * VOL;
    zc *col = (zc *)malloc(sizeof(zc) * (size_t)n);
    zc *rhs = (zc *)malloc(sizeof(zc) * (size_t)n);
    zc *sol = (zc *)malloc(sizeof(zc) * (size_t)n);
    orc_solve_stats st = {0, 0, 0.0};
    int status = 0;
    for (int which = 0; which < 2; ++which) {
        zc *prop = which == 0 ? quark : antiq;
        sysop op = {links, which == 0 ? mass : -mass, which == 0 ? c[0] : conj(c[0]), stepwidth};
        for (int j = 0; j < 6; ++j) {
            for (int64_t x = 1; x <= VOL; ++x)
                for (int d = 0; d < 6; ++d) col[6 * (x - 1) + d] = PRP(prop, x)[d + 6 * j];
            orc_apply_system(links, col, rhs, op.mass, op.c1, stepwidth, 0);
            int64_t its = 0;
            double rr = 0;
            int s = gmres(&op, rhs, sol, n, tol, restart, maxit, &its, &rr);
            if (s) status = -4;
            st.iterations += its;
            if (znrm(rhs, n) > 0) st.solves += 1;
            if (rr > st.max_true_relres) st.max_true_relres = rr;
            for (int64_t x = 1; x <= VOL; ++x)
                for (int d = 0; d < 6; ++d) PRP(prop, x)[d + 6 * j] = sol[6 * (x - 1) + d];
        }
    }
    free(col); free(rhs); free(sol);
    if (stats) *stats = st;
    return status;
}

/* GetNorm_Propagator, nrqcd.f90:573-611: sqrt(sum_x Tr Re(P P^dagger)) */
double orc_prop_norm(const zc *prop)
{
    double s = 0;
    for (int64_t x = 1; x <= VOL; ++x) {
        const zc *p = PRP(prop, x);
        double t = 0;
        for (int k = 0; k < 36; ++k) t += creal(p[k]) * creal(p[k]) + cimag(p[k]) * cimag(p[k]);
        s += t;
    }
    return sqrt(s);
}

static void mm6(const zc *A, const zc *B, zc *C)
{
    zc t[36];
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
            zc s = 0;
            for (int k = 0; k < 6; ++k) s += A[r + 6 * k] * B[k + 6 * c];
            t[r + 6 * c] = s;
        }
    memcpy(C, t, sizeof(t));
}

static void dag6(const zc *A, zc *B)
{
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) B[r + 6 * c] = conj(A[c + 6 * r]);
}

/* Pauli matrices without 1/2 in the reference's convention, su2.f90:68-88 */
static void pauli(int k, zc s[4])
{
    s[0] = s[1] = s[2] = s[3] = 0;
    if (k == 1) { s[1] = 1; s[2] = 1; }
    if (k == 2) { s[1] = -I; s[2] = +I; }
    if (k == 3) { s[0] = 1; s[3] = -1; }
}

void orc_pauli(int k, zc s[4]) { pauli(k, s); }

/* sum_x Tr( i P M Q M ) * w over a list of 6x6 vertex matrices M, with
 * P = A^dag, or P = conj(A) for the _gconjg variants (nrqcd.f90:1728-1905) */
static zc corr_generic(const zc *quark, const zc *antiq, const zc *Ms, int nm, double w, int gconjg)
{
    zc tot = 0;
    for (int64_t x = 1; x <= VOL; ++x) {
        zc ad[36], t1[36], t2[36];
        if (gconjg)
            for (int e = 0; e < 36; ++e) ad[e] = conj(PRP(antiq, x)[e]);
        else
            dag6(PRP(antiq, x), ad);
        zc loc = 0;
        for (int m = 0; m < nm; ++m) {
            const zc *M = Ms + 36 * m;
            mm6(ad, M, t1);
            mm6(t1, PRP(quark, x), t2);
            mm6(t2, M, t1);
            zc tr = 0;
            for (int d = 0; d < 6; ++d) tr += t1[d + 6 * d];
            loc += I * tr * w;
        }
        tot += loc;
    }
    return tot;
}

/* GetMesoncorrelator_O1_1S0_ZeroMomentum, nrqcd.f90:1484-1517 (gconjg: :1728-1761) */
static zc corr_o1_1s0(const zc *quark, const zc *antiq, int gconjg)
{
    zc M[36] = {0};
    for (int d = 0; d < 6; ++d) M[d + 6 * d] = 1;
    return corr_generic(quark, antiq, M, 1, 1.0, gconjg);
}

/* GetMesoncorrelator_O1_3S1_ZeroMomentum, nrqcd.f90:1525-1566: S2CS(sigma_k) (gconjg: :1769-1809) */
static zc corr_o1_3s1(const zc *quark, const zc *antiq, int gconjg)
{
    zc one3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, s[4], Ms[3 * 36];
    for (int k = 1; k <= 3; ++k) {
        pauli(k, s);
        orc_kron(one3, 3, 3, s, 2, 2, Ms + 36 * (k - 1));
    }
    return corr_generic(quark, antiq, Ms, 3, 1.0 / 3.0, gconjg);
}

/* GetMesoncorrelator_O8_1S0_ZeroMomentum, nrqcd.f90:1574-1611: C2CS(T^a) (gconjg: :1817-1857) */
static zc corr_o8_1s0(const zc *quark, const zc *antiq, int gconjg)
{
    init_generators();
    zc one2[4] = {1, 0, 0, 1}, Ms[8 * 36];
    for (int a = 0; a < 8; ++a) orc_kron(GEN[a], 3, 3, one2, 2, 2, Ms + 36 * a);
    return corr_generic(quark, antiq, Ms, 8, 6.0 / 8.0, gconjg);
}

/* GetMesoncorrelator_O8_3S1_ZeroMomentum, nrqcd.f90:1619-1661: T^a (x) sigma_k (gconjg: :1865-1905) */
static zc corr_o8_3s1(const zc *quark, const zc *antiq, int gconjg)
{
    init_generators();
    zc s[4], Ms[24 * 36];
    int m = 0;
    for (int k = 1; k <= 3; ++k) {
        pauli(k, s);
        for (int a = 0; a < 8; ++a) orc_kron(GEN[a], 3, 3, s, 2, 2, Ms + 36 * (m++));
    }
    return corr_generic(quark, antiq, Ms, 24, 6.0 / 24.0, gconjg);
}

zc orc_corr_o1_1s0(const zc *quark, const zc *antiq) { return corr_o1_1s0(quark, antiq, 0); }
zc orc_corr_o1_3s1(const zc *quark, const zc *antiq) { return corr_o1_3s1(quark, antiq, 0); }
zc orc_corr_o8_1s0(const zc *quark, const zc *antiq) { return corr_o8_1s0(quark, antiq, 0); }
zc orc_corr_o8_3s1(const zc *quark, const zc *antiq) { return corr_o8_3s1(quark, antiq, 0); }
zc orc_corr_o1_1s0_gconjg(const zc *quark, const zc *antiq) { return corr_o1_1s0(quark, antiq, 1); }
zc orc_corr_o1_3s1_gconjg(const zc *quark, const zc *antiq) { return corr_o1_3s1(quark, antiq, 1); }
zc orc_corr_o8_1s0_gconjg(const zc *quark, const zc *antiq) { return corr_o8_1s0(quark, antiq, 1); }
zc orc_corr_o8_3s1_gconjg(const zc *quark, const zc *antiq) { return corr_o8_3s1(quark, antiq, 1); }

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
