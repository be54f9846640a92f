"""The C-ABI boundary on the CPU: the library loads, exports every function
declared in include/*.h, and its host-only entry points validate arguments
(no device work is issued here)."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(nrq_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_the_boundary():
    names = _declared()
    for required in ("nrq_lattice_init", "nrq_gauge_leapfrog", "nrq_gauge_box_ic", "nrq_gauge_energy",
                     "nrq_heavy_update", "nrq_cn_init", "nrq_heavy_norms", "nrq_heavy_correlators"):
        assert required in names


def test_library_exports_every_declared_symbol(nrq):
    lib = nrq.library()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_exported_symbols_have_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", str(ROOT / "nrqcd_amd" / "libnrqcd_amd.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (nrq_[a-z0-9_]+)\b", out))
    assert set(_declared()) <= exported


def test_error_message_roundtrip(nrq):
    with pytest.raises(nrq.NRQError) as e:
        nrq.GetLatticePosition(1)
    buf = ctypes.create_string_buffer(256)
    code = nrq.library().nrq_last_error(buf, 256)
    assert code == e.value.code and b"lattice" in buf.value


def test_gauge_calls_reject_null_handles(nrq, lattice):
    lattice((4, 4, 4), (0.05, 1, 1, 1))
    lib = nrq.library()
    assert lib.nrq_gauge_leapfrog(None, 1.0, 1) == -1
    assert lib.nrq_gauge_energy(None, None) == -1
    assert lib.nrq_heavy_update(None, None, 1.0, None, 1.0, None) != 0
    assert lib.nrq_gauge_destroy(None) == 0


def test_fortran_driver_built():
    exe = ROOT / "nrqcd_amd" / "fortran" / "nrqcd_run"
    if not exe.exists():
        pytest.skip("Fortran driver not built (amdflang missing)")
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stdout


def test_cn_shape_setter_validates_and_round_trips(nrq):
    """nrq_cn_set_shape / nrq_cn_get_shape are host-only: the default is 2
    FP64 and 4 FP32 fields per thread, the banded wave-item tiling (4) and the
    XCD remap 8; out-of-range values are rejected and leave the shape as it
    was."""
    default = nrq.cn_get_shape()
    assert default == (2, 4, 4, 8)
    try:
        for shape in ((1, 2, 0, 1), (4, 1, 3, 8), (2, 4, 4, 1)):
            nrq.cn_set_shape(*shape)
            assert nrq.cn_get_shape() == shape
        for bad in ((3, 4, 4, 8), (2, 8, 4, 8), (2, 4, 5, 8), (2, 4, -1, 8), (2, 4, 4, 2)):
            with pytest.raises(nrq.NRQError):
                nrq.cn_set_shape(*bad)
            assert nrq.cn_get_shape() == (2, 4, 4, 1)
    finally:
        nrq.cn_set_shape(*default)


def test_leapfrog_shape_table(nrq, lattice):
    """The autotuner's shape table: the default shape applies everywhere; the
    LDS tiles only to FP32 storage on extents that are multiples of 16, 8 and
    their run of z planes; pinning a shape outside the table is rejected."""
    lattice((32, 16, 24), (0.05, 1, 1, 1))
    shapes = nrq.leapfrog_shapes()
    assert len(shapes) == 19
    assert all(s["applies_fp32"] for s in shapes if not s["tiling"].startswith("LDS"))
    tiles = [s for s in shapes if s["tiling"].startswith("LDS")]
    assert tiles and not any(s["applies"] for s in tiles)
    assert all(s["applies_fp32"] == s["tiling"].endswith("8-plane runs") for s in tiles)
    lattice((16, 6, 5), (0.05, 1, 1, 1))
    assert not any(s["applies_fp32"] for s in nrq.leapfrog_shapes() if s["tiling"].startswith("LDS"))
    with pytest.raises(nrq.NRQError):
        nrq.leapfrog_set_shape(len(shapes))
    nrq.leapfrog_set_shape(-1)
