"""The allocation-free correlator entry point against nrq_heavy_correlators.

nrq_heavy_correlators_noalloc runs the same trace and fold kernels with the
partial sums in the library's fixed device array: on the same field it must
return the same eight numbers bit for bit.  Two calls in a row, the second on
a smaller lattice (the array keeps the first call's partials), the first on a
lattice of more than 4096 x 256 sites, where the trace kernel strides; the
first call comes before anything in the process has reserved the solver
workspace when the file runs alone.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SP = (0.05, 1.0, 1.0, 1.0)


def _noalloc(nrq, h):
    out = np.zeros(4, np.complex128)
    nrq._check(nrq.library().nrq_heavy_correlators_noalloc(h._h, out.ctypes.data_as(ctypes.c_void_p)))
    return out


def _random_field(nrq, v, seed):
    rng = np.random.default_rng(seed)
    shape = (6, 6, v)
    q = np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    a = np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    h = nrq.NRQCDField()
    h.upload(q, a)
    return h, q, a


def test_noalloc_correlators_equal_the_allocating_entry_point(nrq, oracle, lattice):
    for ext, seed in (((104, 104, 100), 11), ((6, 5, 4), 12)):
        print(f"noalloc correlators on {ext}", flush=True)
        lattice(ext, SP)
        v = int(np.prod(ext))
        h, q, a = _random_field(nrq, v, seed)
        got = _noalloc(nrq, h)
        ref = h.correlators()
        want = np.array([ref[k] for k in ("O1_1S0", "O1_3S1", "O8_1S0", "O8_3S1")])
        assert np.array_equal(got, want), (ext, got, want)
        # a second call on the same field (the array keeps the last contents)
        assert np.array_equal(_noalloc(nrq, h), want)
        if v < 1000:
            oracle.lattice_init(ext, SP)
            oc = oracle.correlators(q, a, False)
            for k, x in zip(("O1_1S0", "O1_3S1", "O8_1S0", "O8_3S1"), got):
                assert abs(x - oc[k]) < 1e-12 * abs(oc[k]), k
        h.Destructor()


def test_noalloc_correlators_reject_bad_arguments(nrq, lattice):
    lattice((4, 4, 4), SP)
    lib = nrq.library()
    out = np.zeros(8)
    assert lib.nrq_heavy_correlators_noalloc(None, out.ctypes.data_as(ctypes.c_void_p)) != 0
    h = nrq.NRQCDField()
    assert lib.nrq_heavy_correlators_noalloc(h._h, None) != 0
    lattice((4, 4, 6), SP)
    # a field of the previous lattice
    assert lib.nrq_heavy_correlators_noalloc(h._h, out.ctypes.data_as(ctypes.c_void_p)) != 0
