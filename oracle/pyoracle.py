"""ctypes view of the CPU oracle (liboracle.so) and of oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  Arrays use the reference's
Fortran layouts (links (3,3,3,V) complex128, efield (8,3,V) float64,
propagators (6,6,V) complex128) as Fortran-ordered numpy arrays.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libnrqcd_ref.so"

P = ctypes.c_void_p
I64 = ctypes.c_int64
D = ctypes.c_double
I = ctypes.c_int


class _Z(ctypes.Structure):
    """C99 double _Complex passed or returned by value (SysV: two SSE eightbytes)."""
    _fields_ = [("re", D), ("im", D)]


class OrcSolveStats(ctypes.Structure):
    _fields_ = [("iterations", I64), ("solves", I64), ("max_true_relres", D)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(P)


class Oracle:
    """FP64 CPU restatement of the reference path (see nrqcd_oracle.c)."""

    def __init__(self, path: Path = LIB):
        if not Path(path).exists():
            build()
        self.lib = ctypes.CDLL(str(path))
        L = self.lib
        for name, res, args in [
            ("orc_lattice_init", I, [P, P]), ("orc_lattice_size", I64, []), ("orc_index", I64, [P]),
            ("orc_position", None, [I64, P]), ("orc_neib", I64, [I, I64]), ("orc_negative_index", I64, [I64]),
            ("orc_momentum", None, [I64, P]), ("orc_norm2_momentum", D, [I64]), ("orc_polarisation", None, [P, P]),
            ("orc_generators", None, [P]), ("orc_algebra_matrix", None, [P, P]),
            ("orc_algebra_coordinate", D, [I, P]), ("orc_eigh3", None, [P, P, P]), ("orc_group_exp", None, [P, P]),
            ("orc_kron", None, [P, I, I, P, I, I, P]), ("orc_pauli", None, [I, P]),
            ("orc_ranlxd_init", I, [I, I64]), ("orc_ranlxd", None, [P, I64]), ("orc_ranlxd_get", None, [P]),
            ("orc_normal_cmplx", None, [P, I64]), ("orc_x2p", None, [P]), ("orc_p2x", None, [P]),
            ("orc_cold_init", None, [P, P]), ("orc_box_init", I, [D, D, D, P, P]),
            ("orc_update_links", None, [P, P, D]), ("orc_update_efield", None, [P, P, D]),
            ("orc_leapfrog", None, [P, P, D, I64]), ("orc_energy", D, [P, P]), ("orc_gauss_ref", D, [P, P]),
            ("orc_gauss_max", D, [P, P]), ("orc_plaquette", None, [P, I, I, I64, P]),
            ("orc_mean_plaquette", D, [P]), ("orc_heavy_init_point", None, [P, I64]),
            ("orc_apply_system", None, [P, P, P, D, _Z, D, I]),
            ("orc_cn_update", I, [P, P, P, D, P, D, D, I, I64, P]), ("orc_prop_norm", D, [P]),
            ("orc_num_threads", I, []), ("orc_set_plant", None, [I]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        for name in ["orc_corr_o1_1s0", "orc_corr_o1_3s1", "orc_corr_o8_1s0", "orc_corr_o8_3s1",
                     "orc_corr_o1_1s0_gconjg", "orc_corr_o1_3s1_gconjg", "orc_corr_o8_1s0_gconjg",
                     "orc_corr_o8_3s1_gconjg"]:
            f = getattr(L, name)
            f.restype = _Z
            f.argtypes = [P, P]
        self.vol = 0

    # lattice
    def lattice_init(self, ext, spacing) -> None:
        e = np.ascontiguousarray(ext, np.int64)
        s = np.ascontiguousarray(spacing, np.float64)
        assert self.lib.orc_lattice_init(_ptr(e), _ptr(s)) == 0
        self.vol = int(np.prod(e))
        self.ext = tuple(int(x) for x in e)
        self.spacing = tuple(float(x) for x in s)

    def index(self, pos) -> int:
        return self.lib.orc_index(_ptr(np.ascontiguousarray(pos, np.int64)))

    def position(self, idx) -> np.ndarray:
        p = np.zeros(3, np.int64)
        self.lib.orc_position(idx, _ptr(p))
        return p

    def neib(self, d, idx) -> int:
        return self.lib.orc_neib(d, idx)

    def negative_index(self, idx) -> int:
        return self.lib.orc_negative_index(idx)

    def momentum(self, idx) -> np.ndarray:
        p = np.zeros(3, np.complex128)
        self.lib.orc_momentum(idx, _ptr(p))
        return p

    def norm2_momentum(self, idx) -> float:
        return self.lib.orc_norm2_momentum(idx)

    def polarisation(self, p) -> np.ndarray:
        pp = np.ascontiguousarray(p, np.complex128)
        eps = np.zeros((3, 3), np.complex128, order="F")
        self.lib.orc_polarisation(_ptr(pp), _ptr(eps))
        return eps

    # algebra
    def generators(self) -> np.ndarray:
        t = np.zeros((3, 3, 8), np.complex128, order="F")
        self.lib.orc_generators(_ptr(t))
        return t

    def algebra_matrix(self, alpha) -> np.ndarray:
        a = np.ascontiguousarray(alpha, np.float64)
        m = np.zeros((3, 3), np.complex128, order="F")
        self.lib.orc_algebra_matrix(_ptr(a), _ptr(m))
        return m

    def algebra_coordinate(self, a: int, m) -> float:
        mm = np.asfortranarray(m, np.complex128)
        return self.lib.orc_algebra_coordinate(a, _ptr(mm))

    def group_exp(self, alpha) -> np.ndarray:
        a = np.ascontiguousarray(alpha, np.float64)
        u = np.zeros((3, 3), np.complex128, order="F")
        self.lib.orc_group_exp(_ptr(a), _ptr(u))
        return u

    def kron(self, A, B) -> np.ndarray:
        A = np.asfortranarray(A, np.complex128)
        B = np.asfortranarray(B, np.complex128)
        out = np.zeros((A.shape[0] * B.shape[0], A.shape[1] * B.shape[1]), np.complex128, order="F")
        self.lib.orc_kron(_ptr(A), A.shape[0], A.shape[1], _ptr(B), B.shape[0], B.shape[1], _ptr(out))
        return out

    def pauli(self, k: int) -> np.ndarray:
        s = np.zeros((2, 2), np.complex128, order="F")
        self.lib.orc_pauli(k, _ptr(s))
        return s

    # rng
    def ranlxd_init(self, level: int, seed: int) -> None:
        assert self.lib.orc_ranlxd_init(level, seed) == 0

    def ranlxd(self, n: int) -> np.ndarray:
        r = np.zeros(n)
        self.lib.orc_ranlxd(_ptr(r), n)
        return r

    def ranlxd_get(self) -> np.ndarray:
        s = np.zeros(25, np.int64)
        self.lib.orc_ranlxd_get(_ptr(s))
        return s

    def normal_cmplx(self, n: int) -> np.ndarray:
        z = np.zeros(n, np.complex128)
        self.lib.orc_normal_cmplx(_ptr(z), n)
        return z

    def x2p(self, data) -> np.ndarray:
        d = np.array(data, np.complex128, copy=True)
        self.lib.orc_x2p(_ptr(d))
        return d

    def p2x(self, data) -> np.ndarray:
        d = np.array(data, np.complex128, copy=True)
        self.lib.orc_p2x(_ptr(d))
        return d

    # gauge
    def empty_gauge(self):
        return (np.zeros((3, 3, 3, self.vol), np.complex128, order="F"),
                np.zeros((8, 3, self.vol), np.float64, order="F"))

    def cold(self):
        U, E = self.empty_gauge()
        self.lib.orc_cold_init(_ptr(U), _ptr(E))
        return U, E

    def box_init(self, seed: int, qs: float, amp: float, g: float):
        self.ranlxd_init(2, seed)
        U, E = self.empty_gauge()
        assert self.lib.orc_box_init(qs, amp, g, _ptr(U), _ptr(E)) == 0
        return U, E

    def leapfrog(self, U, E, stepwidth: float = 1.0, nsteps: int = 1) -> None:
        assert U.flags.f_contiguous and E.flags.f_contiguous
        self.lib.orc_leapfrog(_ptr(U), _ptr(E), stepwidth, nsteps)

    def update_links(self, U, E, stepwidth: float) -> None:
        self.lib.orc_update_links(_ptr(U), _ptr(E), stepwidth)

    def update_efield(self, U, E, stepwidth: float) -> None:
        self.lib.orc_update_efield(_ptr(U), _ptr(E), stepwidth)

    def energy(self, U, E) -> float:
        return self.lib.orc_energy(_ptr(U), _ptr(E))

    def gauss_ref(self, U, E) -> float:
        return self.lib.orc_gauss_ref(_ptr(U), _ptr(E))

    def gauss_max(self, U, E) -> float:
        return self.lib.orc_gauss_max(_ptr(U), _ptr(E))

    def mean_plaquette(self, U) -> float:
        return self.lib.orc_mean_plaquette(_ptr(U))

    def plaquette(self, U, i, j, idx) -> np.ndarray:
        P_ = np.zeros((3, 3), np.complex128, order="F")
        self.lib.orc_plaquette(_ptr(U), i, j, idx, _ptr(P_))
        return P_

    # nrqcd
    def point_prop(self, idx: int) -> np.ndarray:
        q = np.zeros((6, 6, self.vol), np.complex128, order="F")
        self.lib.orc_heavy_init_point(_ptr(q), idx)
        return q

    def apply_system(self, U, col, mass, c1, stepwidth, herm: bool) -> np.ndarray:
        """S (herm=False) or S^dagger applied to one 6V column (site-major)."""
        src = np.ascontiguousarray(col, np.complex128)
        out = np.zeros_like(src)
        c = _Z(complex(c1).real, complex(c1).imag)
        self.lib.orc_apply_system(_ptr(U), _ptr(src), _ptr(out), mass, c, stepwidth, int(herm))
        return out

    def cn_update(self, U, Q, A, mass, c, stepwidth=1.0, tol=1e-12, restart=30, maxit=10000):
        cc = np.zeros(4, np.complex128)
        cc[: len(c)] = c
        st = OrcSolveStats()
        rc = self.lib.orc_cn_update(_ptr(U), _ptr(Q), _ptr(A), mass, _ptr(cc), stepwidth, tol, restart, maxit,
                                    ctypes.byref(st))
        return rc, st

    def prop_norm(self, Q) -> float:
        return self.lib.orc_prop_norm(_ptr(Q))

    def correlators(self, Q, A, gconjg=False) -> dict:
        """O1/O8 x 1S0/3S1 at zero momentum; gconjg selects the variants with
        conj(A) in place of A^dag."""
        sfx = "_gconjg" if gconjg else ""
        out = {}
        for key in ("O1_1S0", "O1_3S1", "O8_1S0", "O8_3S1"):
            fn = getattr(self.lib, "orc_corr_" + key.lower() + sfx)
            r = fn(_ptr(Q), _ptr(A))
            out[key] = complex(r.re, r.im)
        return out

    def num_threads(self) -> int:
        return self.lib.orc_num_threads()

    def set_plant(self, which: int) -> None:
        """Plant mistake `which` (nrqcd_oracle.h) into the oracle; 0 removes it.
        Only the tests that show a check can fail use this."""
        self.lib.orc_set_plant(which)


class RefModules:
    """The reference's own pure Fortran modules built by oracle/build_ref.sh."""

    def __init__(self, path: Path = REF_LIB):
        self.lib = ctypes.CDLL(str(path))
        L = self.lib
        for name, res, args in [
            ("ref_ranlxd_init", None, [I64, I64]), ("ref_ranlxd", None, [P, I64]), ("ref_ranlxd_get", None, [P]),
            ("ref_su3_generators", None, [P]), ("ref_su2_generators", None, [P]), ("ref_group_exp", None, [P, P]),
            ("ref_algebra_matrix", None, [P, P]), ("ref_algebra_coordinate", D, [I, P]),
            ("ref_trace_with_generator", None, [I, P, P]), ("ref_expah", None, [P, P]), ("ref_group_log", None, [P, P]),
            ("ref_arglogu", None, [P, P]),
            ("ref_kron", None, [P, I, I, P, I, I, P]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args

    @staticmethod
    def available() -> bool:
        return REF_LIB.exists()

    def ranlxd_init(self, level, seed):
        self.lib.ref_ranlxd_init(level, seed)

    def ranlxd(self, n):
        r = np.zeros(n)
        self.lib.ref_ranlxd(_ptr(r), n)
        return r

    def ranlxd_get(self):
        s = np.zeros(25, np.int64)
        self.lib.ref_ranlxd_get(_ptr(s))
        return s

    def su3_generators(self):
        t = np.zeros((3, 3, 8), np.complex128, order="F")
        self.lib.ref_su3_generators(_ptr(t))
        return t

    def su2_generators(self):
        t = np.zeros((2, 2, 3), np.complex128, order="F")
        self.lib.ref_su2_generators(_ptr(t))
        return t

    def group_exp(self, alpha):
        a = np.ascontiguousarray(alpha, np.float64)
        u = np.zeros((3, 3), np.complex128, order="F")
        self.lib.ref_group_exp(_ptr(a), _ptr(u))
        return u

    def algebra_matrix(self, alpha):
        a = np.ascontiguousarray(alpha, np.float64)
        m = np.zeros((3, 3), np.complex128, order="F")
        self.lib.ref_algebra_matrix(_ptr(a), _ptr(m))
        return m

    def algebra_coordinate(self, a, m):
        mm = np.asfortranarray(m, np.complex128)
        return self.lib.ref_algebra_coordinate(a, _ptr(mm))

    def trace_with_generator(self, a, m):
        mm = np.asfortranarray(m, np.complex128)
        out = np.zeros(1, np.complex128)
        self.lib.ref_trace_with_generator(a, _ptr(mm), _ptr(out))
        return out[0]

    def expah(self, A):
        aa = np.asfortranarray(A, np.complex128)
        out = np.zeros((3, 3), np.complex128, order="F")
        self.lib.ref_expah(_ptr(aa), _ptr(out))
        return out

    def group_log(self, U):
        uu = np.asfortranarray(U, np.complex128)
        out = np.zeros(8)
        self.lib.ref_group_log(_ptr(uu), _ptr(out))
        return out

    def arglogu(self, U):
        uu = np.asfortranarray(U, np.complex128)
        out = np.zeros((3, 3), np.complex128, order="F")
        self.lib.ref_arglogu(_ptr(uu), _ptr(out))
        return out

    def kron(self, A, B):
        A = np.asfortranarray(A, np.complex128)
        B = np.asfortranarray(B, np.complex128)
        out = np.zeros((A.shape[0] * B.shape[0], A.shape[1] * B.shape[1]), np.complex128, order="F")
        self.lib.ref_kron(_ptr(A), A.shape[0], A.shape[1], _ptr(B), B.shape[0], B.shape[1], _ptr(out))
        return out
