"""nrqcd_amd -- MI355X-native real-time SU(3) leapfrog + NRQCD Crank-Nicolson.

Python host layer over the C ABI of ``libnrqcd_amd.so`` (declared in
``include/nrqcd_amd.h``).  The class and method names mirror the reference's
Fortran modules so that scripts read like the reference drivers:

* :mod:`lattice` functions -> :func:`lattice_init` (``lattice::InitModule``,
  lattice.f90:124-228) and the index/neighbour/momentum accessors;
* ``random::InitModule`` -> :func:`random_init` (random.f90:126-140);
* ``type(GaugeConfiguration)`` -> :class:`GaugeConfiguration`
  (gaugeconfiguration_su3.f90:33-116);
* ``nrqcd::InitModule`` -> :func:`nrqcd_init` (nrqcd.f90:167-192) and
  ``type(NRQCDField)`` -> :class:`NRQCDField` (nrqcd.f90:58-117).

Arrays exchanged with the library use the reference's Fortran layouts as
numpy arrays in Fortran order: links ``(3,3,3,V)`` complex128, efield
``(8,3,V)`` float64, propagators ``(6,6,V)`` complex128.

The library needs a HIP device; there is no CPU fallback.  Loading fails
loudly when the shared library is missing.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIBRARY_PATH = _PKG / "libnrqcd_amd.so"

NRQ_OK = 0
NRQ_PRECISION_FP64 = 64
NRQ_PRECISION_FP32 = 32
NRQ_SOLVER_GMRES_FP64 = 0
NRQ_SOLVER_GMRES_MIXED = 1
NRQ_SOLVER_CHEBYSHEV_FP64 = 2
NRQ_SOLVER_CHEBYSHEV_MIXED = 3
NRQ_SOLVER_FGMRES_CNN = 4
NRQ_SOLVER_FGMRES_CNN_MIXED = 5
NRQ_CNN_NWEIGHTS = 4 * 27 * 32 * 32

_ERRORS = {
    -1: "NRQ_ERR_ARG",
    -2: "NRQ_ERR_STATE",
    -3: "NRQ_ERR_HIP",
    -4: "NRQ_ERR_UNSUPPORTED",
    -5: "NRQ_ERR_NOCONV",
    -6: "NRQ_ERR_ALLOC",
}


class NRQError(RuntimeError):
    """Raised when a C-ABI call returns a non-zero status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{_ERRORS.get(code, code)}: {message}")
        self.code = code


class SolveStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int64),
        ("solves", ctypes.c_int64),
        ("restarts", ctypes.c_int64),
        ("max_true_relres", ctypes.c_double),
    ]


_lib = None


def library() -> ctypes.CDLL:
    """Load ``libnrqcd_amd.so`` (built in-tree by ``nrqcd_amd.build``)."""
    global _lib
    if _lib is None:
        path = os.environ.get("NRQCD_AMD_LIB", str(LIBRARY_PATH))
        if not Path(path).exists():
            raise ImportError(
                f"nrqcd_amd: native library {path} is missing; build it with "
                "`python -m nrqcd_amd.build` (there is no CPU fallback)")
        _lib = ctypes.CDLL(path)
        _declare(_lib)
    return _lib


def _declare(lib: ctypes.CDLL) -> None:
    P = ctypes.c_void_p
    I64 = ctypes.c_int64
    D = ctypes.c_double
    I = ctypes.c_int
    sig = {
        "nrq_last_error": (I, [ctypes.c_char_p, ctypes.c_size_t]),
        "nrq_version": (ctypes.c_char_p, []),
        "nrq_event_record": (I, [I]),
        "nrq_event_elapsed": (I, [I, I, P]),
        "nrq_synchronize": (I, []),
        "nrq_bandwidth_probe": (I, [I64, I, P]),
        "nrq_leapfrog_timer_start": (I, [I64]),
        "nrq_leapfrog_timer_stop": (I, [P, P, P, P]),
        "nrq_lattice_init": (I, [P, P]),
        "nrq_lattice_finalize": (I, []),
        "nrq_lattice_size": (I, [P]),
        "nrq_lattice_index": (I, [P, P]),
        "nrq_lattice_position": (I, [I64, P]),
        "nrq_lattice_neib": (I, [I, I64, P]),
        "nrq_lattice_negative_index": (I, [I64, P]),
        "nrq_lattice_momentum": (I, [I64, P]),
        "nrq_lattice_norm2_momentum": (I, [I64, P]),
        "nrq_lattice_max_norm2_momentum": (I, [P]),
        "nrq_random_init": (I, [I64]),
        "nrq_random_uniform": (I, [P, I64]),
        "nrq_gauge_create": (I, [I, P]),
        "nrq_gauge_destroy": (I, [P]),
        "nrq_gauge_copy": (I, [P, P]),
        "nrq_gauge_cold": (I, [P]),
        "nrq_gauge_box_ic": (I, [P, D, D, D]),
        "nrq_gauge_upload": (I, [P, P, P]),
        "nrq_gauge_download": (I, [P, P, P]),
        "nrq_gauge_get_link": (I, [P, I, I64, P]),
        "nrq_gauge_get_efield": (I, [P, I, I64, P]),
        "nrq_gauge_leapfrog": (I, [P, D, I64]),
        "nrq_gauge_leapfrog_batch": (I, [P, I, D, I64]),
        "nrq_gauge_energy": (I, [P, P]),
        "nrq_gauge_gauss_ref": (I, [P, P]),
        "nrq_gauge_gauss_max": (I, [P, P]),
        "nrq_gauge_mean_plaquette": (I, [P, P]),
        "nrq_gauge_observables": (I, [P, P]),
        "nrq_gauge_log_observables": (I, [P]),
        "nrq_gauge_log_read": (I, [P, P, I64, P]),
        "nrq_gauge_plaquette": (I, [P, I, I, I64, P]),
        "nrq_gauge_reunitarise": (I, [P]),
        "nrq_gauge_set_reunitarise_interval": (I, [P, I64]),
        "nrq_gauge_launch_shape": (I, [P, P, P, P, P, P]),
        "nrq_leapfrog_shape_count": (I, [P]),
        "nrq_leapfrog_shape_info": (I, [I, P, P, P, P, P, P]),
        "nrq_leapfrog_set_shape": (I, [I]),
        "nrq_leapfrog_get_shape": (I, [P]),
        "nrq_cn_set_shape": (I, [I, I, I, I]),
        "nrq_cn_get_shape": (I, [P, P, P, P]),
        "nrq_cn_fallbacks": (I, [P]),
        "nrq_gauge_coulomb_gaugefix": (I, [P, D, D, I64, P, P]),
        "nrq_gauge_transverse_correlators": (I, [P, P, P]),
        "nrq_cn_init": (I, [D, P, D, D]),
        "nrq_cn_finalize": (I, []),
        "nrq_cn_set_solver": (I, [I, I, I64]),
        "nrq_heavy_create": (I, [P]),
        "nrq_heavy_destroy": (I, [P]),
        "nrq_heavy_copy": (I, [P, P]),
        "nrq_heavy_init_point": (I, [P, I64, I64]),
        "nrq_heavy_init_point_dof": (I, [P, I, I, I64, I, I, I64]),
        "nrq_heavy_upload": (I, [P, P, P]),
        "nrq_heavy_download": (I, [P, P, P]),
        "nrq_heavy_update": (I, [P, P, D, P, D, P]),
        "nrq_heavy_update_batch": (I, [P, P, I, P, P, D, P]),
        "nrq_cnn_create": (I, [P, ctypes.c_uint64]),
        "nrq_cnn_destroy": (I, [P]),
        "nrq_cnn_get_weights": (I, [P, P]),
        "nrq_cnn_set_weights": (I, [P, P]),
        "nrq_cnn_forward": (I, [P, P, P]),
        "nrq_cnn_time_forward": (I, [P, I, P]),
        "nrq_cnn_loss_grad": (I, [P, P, P, D, P, P]),
        "nrq_cnn_loss_grad_trace": (I, [P, P, P, D, P, P, P, P]),
        "nrq_cnn_harvest": (I, [P, P, D, D, I, ctypes.c_uint64]),
        "nrq_cnn_dataset_size": (I, [P, P]),
        "nrq_cnn_train": (I, [P, I, D, ctypes.c_uint64, P]),
        "nrq_cn_set_preconditioner": (I, [P]),
        "nrq_heavy_norms": (I, [P, P, P]),
        "nrq_heavy_correlators": (I, [P, P]),
        "nrq_heavy_correlators_gconjg": (I, [P, P]),
        "nrq_heavy_correlators_noalloc": (I, [P, P]),
        "nrq_spectral_functions": (I, [P, I64, I64, D, P, P]),
        "nrq_gauge_box_ic_correlators": (I, [P, D, D, D, P, P]),
        "nrq_gauge_transverse_init": (I, [P, P, I, P, P]),
        "nrq_gauge_equilibrium_init": (I, [P, D, I64, I64, D, P, P, P, P]),
        "nrq_t2w": (I, [P, I64, D]),
        "nrq_w2t": (I, [P, I64, D]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _check(rc: int) -> None:
    if rc != NRQ_OK:
        buf = ctypes.create_string_buffer(512)
        library().nrq_last_error(buf, 512)
        raise NRQError(rc, buf.value.decode(errors="replace"))


def event_record(slot: int) -> None:
    """Record HIP event ``slot`` on the library stream."""
    _check(library().nrq_event_record(slot))


def event_elapsed_ms(start: int, stop: int) -> float:
    ms = ctypes.c_double()
    _check(library().nrq_event_elapsed(start, stop, ctypes.byref(ms)))
    return ms.value


def synchronize() -> None:
    _check(library().nrq_synchronize())


def leapfrog_timer_start(max_spans: int = 1024) -> None:
    """Start timing the fused leapfrog launches (one HIP event pair per
    leapfrog call, on the library stream)."""
    _check(library().nrq_leapfrog_timer_start(max_spans))


def leapfrog_timer_stop() -> dict:
    """Stop the fused-launch timer: summed event time (ms), fused launches,
    (site, configuration) steps they hold, and untimed launches."""
    ms, la, ss, sk = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    _check(library().nrq_leapfrog_timer_stop(ctypes.byref(ms), ctypes.byref(la), ctypes.byref(ss), ctypes.byref(sk)))
    return {"ms": ms.value, "launches": la.value, "site_steps": ss.value, "skipped": sk.value}


def bandwidth_probe(nbytes: int = 1 << 30, reps: int = 20) -> float:
    """Read+write GB/s of a plain streaming-copy kernel over ``nbytes``."""
    out = ctypes.c_double()
    _check(library().nrq_bandwidth_probe(nbytes, reps, ctypes.byref(out)))
    return out.value


# ---------------------------------------------------------------------------
# lattice / random
# ---------------------------------------------------------------------------
def lattice_init(extensions, spacings) -> None:
    """``lattice::InitModule(LatticeExtensions_, LatticeSpacings_)``."""
    ext = np.ascontiguousarray(extensions, dtype=np.int64)
    sp = np.ascontiguousarray(spacings, dtype=np.float64)
    if ext.shape != (3,) or sp.shape != (4,):
        raise ValueError("extensions must have 3 entries, spacings 4 (a0..a3)")
    _check(library().nrq_lattice_init(_ptr(ext), _ptr(sp)))


def lattice_finalize() -> None:
    _check(library().nrq_lattice_finalize())


def lattice_size() -> int:
    v = ctypes.c_int64()
    _check(library().nrq_lattice_size(ctypes.byref(v)))
    return v.value


def GetLatticeIndex(position) -> int:
    pos = np.ascontiguousarray(position, dtype=np.int64)
    out = ctypes.c_int64()
    _check(library().nrq_lattice_index(_ptr(pos), ctypes.byref(out)))
    return out.value


def GetLatticePosition(index: int) -> np.ndarray:
    pos = np.zeros(3, np.int64)
    _check(library().nrq_lattice_position(index, _ptr(pos)))
    return pos


def GetNeib_G(direction: int, index: int) -> int:
    out = ctypes.c_int64()
    _check(library().nrq_lattice_neib(direction, index, ctypes.byref(out)))
    return out.value


def GetNegativeLatticeIndex(index: int) -> int:
    out = ctypes.c_int64()
    _check(library().nrq_lattice_negative_index(index, ctypes.byref(out)))
    return out.value


def GetMomentum_G(index: int) -> np.ndarray:
    p = np.zeros(3, np.complex128)
    _check(library().nrq_lattice_momentum(index, _ptr(p)))
    return p


def GetNorm2Momentum_G(index: int) -> float:
    out = ctypes.c_double()
    _check(library().nrq_lattice_norm2_momentum(index, ctypes.byref(out)))
    return out.value


def GetMaxNorm2Momentum() -> float:
    out = ctypes.c_double()
    _check(library().nrq_lattice_max_norm2_momentum(ctypes.byref(out)))
    return out.value


def momentum_norms() -> np.ndarray:
    """|p| of every momentum index (lattice order), as written next to the
    gluon distribution (main.f90:3093-3189)."""
    n = lattice_size()
    return np.array([GetNorm2Momentum_G(k) for k in range(1, n + 1)])


def random_init(seed: int) -> None:
    """``random::InitModule(CustomSeed)`` -> ranlxd level 2."""
    _check(library().nrq_random_init(seed))


def random_uniform(n: int) -> np.ndarray:
    r = np.zeros(n)
    _check(library().nrq_random_uniform(_ptr(r), n))
    return r


# ---------------------------------------------------------------------------
# gaugeconfiguration_su3
# ---------------------------------------------------------------------------
class GaugeConfiguration:
    """Device-resident ``type(GaugeConfiguration)``.

    ``precision`` selects FP64 (reference) or FP32 link/E storage."""

    def __init__(self, precision: int = NRQ_PRECISION_FP64):
        self._h = ctypes.c_void_p()
        _check(library().nrq_gauge_create(precision, ctypes.byref(self._h)))
        self.precision = precision

    # -- lifetime --------------------------------------------------------
    def Deallocate(self) -> None:
        if self._h:
            library().nrq_gauge_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.Deallocate()
        except Exception:
            pass

    def copy(self) -> "GaugeConfiguration":
        """Deep copy (the reference's intrinsic assignment)."""
        other = GaugeConfiguration.__new__(GaugeConfiguration)
        other._h = ctypes.c_void_p()
        _check(library().nrq_gauge_create(self.precision, ctypes.byref(other._h)))
        other.precision = self.precision
        _check(library().nrq_gauge_copy(other._h, self._h))
        return other

    # -- initialisation --------------------------------------------------
    def ColdInit(self) -> None:
        _check(library().nrq_gauge_cold(self._h))

    def TransversePolarisedOccupiedInit_Box(self, SaturationScale: float, Amplitude: float,
                                            Coupling: float, correlators: bool = False):
        """gaugeconfiguration_su3.f90:906-951; with correlators, also returns the
        momentum-space (aa, ee) per momentum index (its optional outputs)."""
        if not correlators:
            _check(library().nrq_gauge_box_ic(self._h, SaturationScale, Amplitude, Coupling))
            return None
        v = lattice_size()
        aa, ee = np.zeros(v), np.zeros(v)
        _check(library().nrq_gauge_box_ic_correlators(self._h, SaturationScale, Amplitude, Coupling, _ptr(aa),
                                                      _ptr(ee)))
        return aa, ee

    def _transverse_init(self, Occupation, sender: int, correlators: bool):
        v = lattice_size()
        occ = np.ascontiguousarray(Occupation, np.float64)
        if occ.size != v:
            raise ValueError(f"occupation has {occ.size} entries, the lattice {v}")
        aa = np.zeros(v) if correlators else None
        ee = np.zeros(v) if correlators else None
        _check(library().nrq_gauge_transverse_init(self._h, _ptr(occ), sender, _ptr(aa), _ptr(ee)))
        return (aa, ee) if correlators else None

    def EquilibriumInit(self, Beta: float, nefieldinit: int, nequilibrium: int, tolerance_Eprojection=None,
                        ChargeDensity=None):
        """gaugeconfiguration_su3.f90:380-814; returns per outer iteration
        (energy, Gauss deviation after the leapfrog steps, projection sweeps)."""
        v = lattice_size()
        rho = None
        if ChargeDensity is not None:
            rho = np.asfortranarray(ChargeDensity, np.float64).reshape((3, v), order="F")
        en = np.zeros(max(nefieldinit, 1))
        dev = np.zeros(max(nefieldinit, 1))
        nproj = np.zeros(max(nefieldinit, 1), np.int64)
        tol = -1.0 if tolerance_Eprojection is None else float(tolerance_Eprojection)
        _check(library().nrq_gauge_equilibrium_init(self._h, float(Beta), int(nefieldinit), int(nequilibrium), tol,
                                                    _ptr(rho), _ptr(en), _ptr(dev), _ptr(nproj)))
        return en[:nefieldinit], dev[:nefieldinit], nproj[:nefieldinit]

    def TransversePolarisedOccupiedInit(self, Occupation, correlators: bool = False):
        """gaugeconfiguration_su3.f90:1263-1446 for any occupation by lattice index."""
        return self._transverse_init(Occupation, 0, correlators)

    def TransversePolarisedOccupiedInit_specificSender(self, Occupation, correlators: bool = False):
        """gaugeconfiguration_su3.f90:959-1255: amplitudes drawn in momentum space."""
        return self._transverse_init(Occupation, 1, correlators)

    # -- state exchange ----------------------------------------------------
    def upload(self, links: np.ndarray | None = None, efield: np.ndarray | None = None) -> None:
        v = lattice_size()
        lk = None if links is None else np.asfortranarray(links, dtype=np.complex128).reshape((3, 3, 3, v), order="F")
        ef = None if efield is None else np.asfortranarray(efield, dtype=np.float64).reshape((8, 3, v), order="F")
        _check(library().nrq_gauge_upload(self._h, _ptr(lk), _ptr(ef)))

    def download(self):
        v = lattice_size()
        lk = np.zeros((3, 3, 3, v), np.complex128, order="F")
        ef = np.zeros((8, 3, v), np.float64, order="F")
        _check(library().nrq_gauge_download(self._h, _ptr(lk), _ptr(ef)))
        return lk, ef

    def GetLink_G(self, i: int, LatticeIndex: int) -> np.ndarray:
        out = np.zeros((3, 3), np.complex128, order="F")
        _check(library().nrq_gauge_get_link(self._h, i, LatticeIndex, _ptr(out)))
        return out

    def GetEfield_G(self, i: int, LatticeIndex: int) -> np.ndarray:
        out = np.zeros(8)
        _check(library().nrq_gauge_get_efield(self._h, i, LatticeIndex, _ptr(out)))
        return out

    # -- evolution -------------------------------------------------------
    def Update(self, StepWidth: float = 1.0, nsteps: int = 1) -> None:
        """``GaugeConf%Update([StepWidth])`` applied ``nsteps`` times."""
        _check(library().nrq_gauge_leapfrog(self._h, StepWidth, nsteps))

    def Reunitarise(self) -> None:
        _check(library().nrq_gauge_reunitarise(self._h))

    def SetReunitarisationInterval(self, interval: int) -> None:
        """Reunitarise (FP64) after every `interval` leapfrog steps; 0: never."""
        _check(library().nrq_gauge_set_reunitarise_interval(self._h, int(interval)))

    def launch_shape(self) -> dict:
        """Launch shape chosen by the leapfrog autotuner."""
        bs, xcd, tiling, nt, rec = (ctypes.c_int() for _ in range(5))
        _check(library().nrq_gauge_launch_shape(self._h, ctypes.byref(bs), ctypes.byref(xcd), ctypes.byref(tiling),
                                                ctypes.byref(nt), ctypes.byref(rec)))
        return {"threads_per_block": bs.value, "xcd_remap": xcd.value, "tiling": _TILINGS[tiling.value],
                "streaming": bool(nt.value), "two_row_neighbours": bool(rec.value)}

    # -- observables -----------------------------------------------------
    def observables(self) -> dict:
        out = np.zeros(4)
        _check(library().nrq_gauge_observables(self._h, _ptr(out)))
        return {"energy": out[0], "gauss_ref": out[1], "gauss_max": out[2], "plaquette": out[3]}

    def log_observables(self) -> None:
        """Append the observables of the current configuration to the
        device-side log without waiting for the GPU."""
        _check(library().nrq_gauge_log_observables(self._h))

    def read_log(self, max_rows: int | None = None) -> np.ndarray:
        """Oldest logged rows (energy, gauss_ref, gauss_max, plaquette),
        removed from the log (all of them by default); waits for the GPU."""
        parts, left = [], max_rows
        while left is None or left > 0:
            cap = 4096 if left is None else min(int(left), 4096)
            out = np.zeros((cap, 4))
            rows = ctypes.c_int64()
            _check(library().nrq_gauge_log_read(self._h, _ptr(out), cap, ctypes.byref(rows)))
            parts.append(out[: rows.value])
            if left is not None:
                left -= rows.value
            if rows.value < cap:
                break
        return np.concatenate(parts) if parts else np.zeros((0, 4))

    def GetEnergy(self) -> float:
        return self.observables()["energy"]

    def GetDeviationFromGaussLaw(self) -> float:
        return self.observables()["gauss_ref"]

    def GetGaussMax(self) -> float:
        return self.observables()["gauss_max"]

    def GetMeanPlaquette(self) -> float:
        return self.observables()["plaquette"]

    def GetPlaquette_G(self, i: int, j: int, LatticeIndex: int) -> np.ndarray:
        out = np.zeros((3, 3), np.complex128, order="F")
        _check(library().nrq_gauge_plaquette(self._h, i, j, LatticeIndex, _ptr(out)))
        return out

    # -- gluon occupation --------------------------------------------------
    def CoulombGaugefixing(self, Tolerance_: float = 1e-8, alpha_: float = 0.08, MaxIterations_: int = 100):
        """Davies' Fourier-accelerated Coulomb gauge fixing of the links (E is
        not transformed, as in the reference).  Returns (gauge transformations
        applied, final max_x ||Delta(x)||_F / 9)."""
        it = ctypes.c_int64()
        dev = ctypes.c_double()
        _check(library().nrq_gauge_coulomb_gaugefix(self._h, Tolerance_, alpha_, MaxIterations_, ctypes.byref(it),
                                                     ctypes.byref(dev)))
        return it.value, dev.value

    def transverse_correlators(self):
        """(AA, EE) per momentum index (lattice order)."""
        n = lattice_size()
        aa, ee = np.zeros(n), np.zeros(n)
        _check(library().nrq_gauge_transverse_correlators(self._h, _ptr(aa), _ptr(ee)))
        return aa, ee

    def GetTransverseAACorrelator(self) -> np.ndarray:
        n = lattice_size()
        aa = np.zeros(n)
        _check(library().nrq_gauge_transverse_correlators(self._h, _ptr(aa), None))
        return aa

    def GetTransverseEECorrelator(self) -> np.ndarray:
        n = lattice_size()
        ee = np.zeros(n)
        _check(library().nrq_gauge_transverse_correlators(self._h, None, _ptr(ee)))
        return ee

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h


# ---------------------------------------------------------------------------
# nrqcd
# ---------------------------------------------------------------------------
_TILINGS = ("consecutive sites", "z-sweep tiles of x-rows", "z-sweep 16-wide tiles", "LDS tiles, 16-plane runs",
            "LDS tiles, 8-plane runs")


def leapfrog_shapes() -> list:
    """The autotuner's candidate launch shapes, with whether each applies to
    the current lattice with FP64 storage ("applies") and with FP32 storage
    ("applies_fp32")."""
    n = ctypes.c_int()
    _check(library().nrq_leapfrog_shape_count(ctypes.byref(n)))
    out = []
    for k in range(n.value):
        v = [ctypes.c_int() for _ in range(6)]
        _check(library().nrq_leapfrog_shape_info(k, *[ctypes.byref(x) for x in v]))
        out.append({"threads_per_block": v[0].value, "xcd_remap": v[1].value, "tiling": _TILINGS[v[2].value],
                    "streaming": bool(v[3].value), "two_row_neighbours": bool(v[4].value),
                    "applies": bool(v[5].value & 1), "applies_fp32": bool(v[5].value & 2)})
    return out


def leapfrog_set_shape(k: int) -> None:
    """Pin launch shape k for every later leapfrog call (-1: autotuner)."""
    _check(library().nrq_leapfrog_set_shape(int(k)))


def leapfrog_get_shape() -> int:
    k = ctypes.c_int()
    _check(library().nrq_leapfrog_get_shape(ctypes.byref(k)))
    return k.value


def cn_set_shape(nc64: int, nc32: int, tiling: int = 4, xcd_remap: int = 8) -> None:
    """Launch shape of the CN stencil kernels (fields per thread in FP64 and
    FP32, site tiling 0-4 with 3 the wave-item map and 4 its banded z-sweep
    order, XCD remap)."""
    _check(library().nrq_cn_set_shape(nc64, nc32, tiling, xcd_remap))


def cn_get_shape() -> tuple:
    v = [ctypes.c_int() for _ in range(4)]
    _check(library().nrq_cn_get_shape(*[ctypes.byref(x) for x in v]))
    return tuple(x.value for x in v)


def cn_fallbacks() -> int:
    """Column solves the learned preconditioner handed to plain GMRES."""
    n = ctypes.c_int64()
    _check(library().nrq_cn_fallbacks(ctypes.byref(n)))
    return n.value


def leapfrog_batch(configs, StepWidth: float = 1.0, nsteps: int = 1) -> None:
    """Advance an ensemble of GaugeConfiguration objects (one storage
    precision) by ``nsteps`` leapfrog steps, one launch per pass over all of
    them, replayed from a captured hipGraph."""
    hs = (ctypes.c_void_p * len(configs))(*[c.handle.value for c in configs])
    _check(library().nrq_gauge_leapfrog_batch(ctypes.cast(hs, ctypes.c_void_p), len(configs), StepWidth, nsteps))


def update_batch(fields, GaugeConf, masses, WilsonCoeffs, StepWidth: float = 1.0) -> "SolveStats":
    """One CN step of several NRQCDField objects with their own masses on one
    gauge configuration, as one batched multi-right-hand-side solve."""
    if len(fields) != len(masses):
        raise ValueError("one mass per field")
    hs = (ctypes.c_void_p * len(fields))(*[f._h.value for f in fields])
    ms = np.ascontiguousarray(masses, dtype=np.float64)
    c = _coeffs(WilsonCoeffs)
    st = SolveStats()
    _check(library().nrq_heavy_update_batch(ctypes.cast(hs, ctypes.c_void_p), _ptr(ms), len(fields), GaugeConf.handle,
                                            _ptr(c), StepWidth, ctypes.byref(st)))
    return st


def _coeffs(WilsonCoeffs) -> np.ndarray:
    c = np.zeros(4, np.complex128)
    c[: len(WilsonCoeffs)] = WilsonCoeffs
    return c


def nrqcd_init(Mass: float, WilsonCoeffs, StepWidth: float, kspTol: float) -> None:
    """``nrqcd::InitModule(Mass, WilsonCoeffs, StepWidth, kspTol)``."""
    c = _coeffs(WilsonCoeffs)
    _check(library().nrq_cn_init(Mass, _ptr(c), StepWidth, kspTol))


def nrqcd_finalize() -> None:
    _check(library().nrq_cn_finalize())


def nrqcd_set_solver(kind: int = NRQ_SOLVER_GMRES_FP64, restart: int = 30, maxit: int = 10000) -> None:
    _check(library().nrq_cn_set_solver(kind, restart, maxit))


class NRQCDField:
    """Device-resident ``type(NRQCDField)`` (quark and antiquark propagator)."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        _check(library().nrq_heavy_create(ctypes.byref(self._h)))

    def Destructor(self) -> None:
        if self._h:
            library().nrq_heavy_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.Destructor()
        except Exception:
            pass

    def copy(self) -> "NRQCDField":
        other = NRQCDField()
        _check(library().nrq_heavy_copy(other._h, self._h))
        return other

    def InitSinglePoint(self, LatticeIndex_quark: int, LatticeIndex_antiq: int) -> None:
        _check(library().nrq_heavy_init_point(self._h, LatticeIndex_quark, LatticeIndex_antiq))

    def InitSinglePointSingleDoF(self, spin_quark, colour_quark, LatticeIndex_quark, spin_antiq, colour_antiq,
                                 LatticeIndex_antiq) -> None:
        _check(library().nrq_heavy_init_point_dof(self._h, spin_quark, colour_quark, LatticeIndex_quark,
                                                  spin_antiq, colour_antiq, LatticeIndex_antiq))

    def upload(self, QuarkProp=None, AntiQProp=None) -> None:
        v = lattice_size()
        q = None if QuarkProp is None else np.asfortranarray(QuarkProp, dtype=np.complex128).reshape((6, 6, v), order="F")
        a = None if AntiQProp is None else np.asfortranarray(AntiQProp, dtype=np.complex128).reshape((6, 6, v), order="F")
        _check(library().nrq_heavy_upload(self._h, _ptr(q), _ptr(a)))

    def download(self):
        v = lattice_size()
        q = np.zeros((6, 6, v), np.complex128, order="F")
        a = np.zeros((6, 6, v), np.complex128, order="F")
        _check(library().nrq_heavy_download(self._h, _ptr(q), _ptr(a)))
        return q, a

    def Update(self, GaugeConf: GaugeConfiguration, Mass: float, WilsonCoeffs, StepWidth: float = 1.0) -> SolveStats:
        c = _coeffs(WilsonCoeffs)
        st = SolveStats()
        _check(library().nrq_heavy_update(self._h, GaugeConf.handle, Mass, _ptr(c), StepWidth, ctypes.byref(st)))
        return st

    def norms(self):
        q = ctypes.c_double()
        a = ctypes.c_double()
        _check(library().nrq_heavy_norms(self._h, ctypes.byref(q), ctypes.byref(a)))
        return q.value, a.value

    def GetNorm_Quark(self) -> float:
        return self.norms()[0]

    def GetNorm_AntiQ(self) -> float:
        return self.norms()[1]

    def correlators(self, gconjg: bool = False) -> dict:
        """GetMesoncorrelator_{O1,O8}_{1S0,3S1}_ZeroMomentum (nrqcd.f90:1484-1661),
        or their _gconjg variants (:1728-1905) with conj(A) in place of A^dag."""
        out = np.zeros(4, np.complex128)
        fn = library().nrq_heavy_correlators_gconjg if gconjg else library().nrq_heavy_correlators
        _check(fn(self._h, _ptr(out)))
        return {"O1_1S0": out[0], "O1_3S1": out[1], "O8_1S0": out[2], "O8_3S1": out[3]}

    def GetMesoncorrelator_O1_3S1_ZeroMomentum(self) -> complex:
        return self.correlators()["O1_3S1"]

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h


class PreconditionerNet:
    """The learned CN preconditioner (``nrq_cnn``): a 4-layer 3-D CNN on 18
    real channels per site, trained on the GPU with the library's own
    forward/backward kernels and used by ``NRQ_SOLVER_FGMRES_CNN``."""

    def __init__(self, seed: int = 1):
        self._h = ctypes.c_void_p()
        _check(library().nrq_cnn_create(ctypes.byref(self._h), seed))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            library().nrq_cnn_destroy(self._h)
            self._h = ctypes.c_void_p()

    @property
    def weights(self) -> np.ndarray:
        w = np.zeros(NRQ_CNN_NWEIGHTS, np.float32)
        _check(library().nrq_cnn_get_weights(self._h, _ptr(w)))
        return w

    @weights.setter
    def weights(self, w) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != NRQ_CNN_NWEIGHTS:
            raise ValueError("expected %d weights" % NRQ_CNN_NWEIGHTS)
        _check(library().nrq_cnn_set_weights(self._h, _ptr(w)))

    def forward(self, inp: np.ndarray) -> np.ndarray:
        """N(inp) for inp of shape (V, 18) on the current lattice."""
        inp = np.ascontiguousarray(inp, dtype=np.float32)
        out = np.zeros_like(inp)
        _check(library().nrq_cnn_forward(self._h, _ptr(inp), _ptr(out)))
        return out

    def time_forward(self, reps: int = 20) -> float:
        """Mean device milliseconds of one network apply on the current
        lattice (inputs resident on the GPU)."""
        ms = ctypes.c_double()
        _check(library().nrq_cnn_time_forward(self._h, int(reps), ctypes.byref(ms)))
        return ms.value

    def loss_grad(self, inp: np.ndarray, target: np.ndarray, norm: float = 1.0):
        inp = np.ascontiguousarray(inp, dtype=np.float32)
        target = np.ascontiguousarray(target, dtype=np.float32)
        loss = ctypes.c_double()
        grad = np.zeros(NRQ_CNN_NWEIGHTS, np.float32)
        _check(library().nrq_cnn_loss_grad(self._h, _ptr(inp), _ptr(target), norm, ctypes.byref(loss), _ptr(grad)))
        return loss.value, grad

    def loss_grad_trace(self, inp: np.ndarray, target: np.ndarray, norm: float = 1.0):
        """loss_grad with the network's own BF16 layer inputs and output
        gradients: (loss, grad, acts[4][V][32], deltas[4][V][32])."""
        inp = np.ascontiguousarray(inp, dtype=np.float32)
        target = np.ascontiguousarray(target, dtype=np.float32)
        v = inp.shape[0]
        loss = ctypes.c_double()
        grad = np.zeros(NRQ_CNN_NWEIGHTS, np.float32)
        acts = np.zeros((4, v, 32), np.float32)
        deltas = np.zeros((4, v, 32), np.float32)
        _check(library().nrq_cnn_loss_grad_trace(self._h, _ptr(inp), _ptr(target), norm, ctypes.byref(loss), _ptr(grad),
                                                 _ptr(acts), _ptr(deltas)))
        return loss.value, grad, acts, deltas

    def harvest(self, GaugeConf, mass: float, nsolves: int, seed: int = 1, StepWidth: float = 1.0) -> None:
        """Training pairs from plain GMRES solves of CN systems at this mass."""
        _check(library().nrq_cnn_harvest(self._h, GaugeConf.handle, mass, StepWidth, nsolves, seed))

    def dataset_size(self) -> int:
        n = ctypes.c_int64()
        _check(library().nrq_cnn_dataset_size(self._h, ctypes.byref(n)))
        return n.value

    def train(self, epochs: int, lr: float = 1e-3, seed: int = 1) -> float:
        loss = ctypes.c_double()
        _check(library().nrq_cnn_train(self._h, epochs, lr, seed, ctypes.byref(loss)))
        return loss.value

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h


def nrqcd_set_preconditioner(net) -> None:
    """Network used by NRQ_SOLVER_FGMRES_CNN (None: none)."""
    _check(library().nrq_cn_set_preconditioner(net.handle if net is not None else None))


def spectral_functions(signals, dt: float, stages: bool = False):
    """ComputeSpectrum (main.f90:1865-2100) on the GPU for a batch of
    correlators: signals (batch, n) or (n,) complex at t = 0, dt, ...;
    returns spectra (batch, 2n - 2) in transform order (and, with stages,
    the symmetrised, windowed and reordered signals, (batch, 3, 2n - 2))."""
    sig = np.ascontiguousarray(np.atleast_2d(np.asarray(signals, np.complex128)))
    batch, n = sig.shape
    N = max(2 * n - 2, 0)
    spec = np.zeros((batch, N))
    st = np.zeros((batch, 3, N), np.complex128) if stages else None
    _check(library().nrq_spectral_functions(_ptr(sig), n, batch, float(dt), _ptr(spec), _ptr(st)))
    squeeze = np.asarray(signals).ndim == 1
    if squeeze:
        spec = spec[0]
        st = st[0] if stages else None
    return (spec, st) if stages else spec


def t2w(data, dt: float) -> np.ndarray:
    """twfft.f90:73-105: dt sum_j f_j exp(+2 pi i j l / n)"""
    x = np.ascontiguousarray(np.asarray(data, np.complex128)).copy()
    _check(library().nrq_t2w(_ptr(x), x.size, float(dt)))
    return x


def w2t(data, dt: float) -> np.ndarray:
    """twfft.f90:33-66: sum_l F_l exp(-2 pi i j l / n) / (dt n)"""
    x = np.ascontiguousarray(np.asarray(data, np.complex128)).copy()
    _check(library().nrq_w2t(_ptr(x), x.size, float(dt)))
    return x
