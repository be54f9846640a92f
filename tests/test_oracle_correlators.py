"""Pinning the oracle's zero-momentum meson correlators (O1/O8 x 1S0/3S1 and
their _gconjg variants, nrqcd.f90:1484-1661 and 1728-1905) on general
propagators; no GPU.

The vertex matrices are built here from the reference's compiled generator
and Pauli tables (tests/golden/ref_vectors.npz) with numpy's Kronecker
product, which the same fixtures show to equal the reference's GetKronProd.
The oracle builds them with its own tables and Kronecker routine, so a
transposed or mis-ordered vertex, a wrong weight or a dropped conjugation
fails here; the free closed form only sees propagators proportional to the
identity."""
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden" / "ref_vectors.npz"
KEYS = ("O1_1S0", "O1_3S1", "O8_1S0", "O8_3S1")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLDEN, allow_pickle=False)


def _vertices(gold):
    T = gold["su3_generators"]
    S = gold["su2_generators"]
    # equal up to the rounding of the complex products A_ij B_kl
    assert np.allclose(np.kron(gold["kron_A"], gold["kron_B"]), gold["kron_AB"], rtol=0, atol=1e-15)
    one2, one3 = np.eye(2), np.eye(3)
    return {
        "O1_1S0": ([np.eye(6)], 1.0),
        "O1_3S1": ([np.kron(one3, S[:, :, k]) for k in range(3)], 1 / 3),
        "O8_1S0": ([np.kron(T[:, :, a], one2) for a in range(8)], 6 / 8),
        "O8_3S1": ([np.kron(T[:, :, a], S[:, :, k]) for k in range(3) for a in range(8)], 6 / 24),
    }


def _random_props(oracle, rng, ext):
    oracle.lattice_init(ext, (0.05, 1.0, 1.0, 1.0))
    vol = int(np.prod(ext))
    Q = np.asfortranarray(rng.normal(size=(6, 6, vol)) + 1j * rng.normal(size=(6, 6, vol)))
    A = np.asfortranarray(rng.normal(size=(6, 6, vol)) + 1j * rng.normal(size=(6, 6, vol)))
    return Q, A


@pytest.mark.parametrize("gconjg", [False, True])
def test_correlators_match_reference_vertices(oracle, gold, gconjg):
    rng = np.random.default_rng(41 + gconjg)
    Q, A = _random_props(oracle, rng, (2, 3, 2))
    got = oracle.correlators(Q, A, gconjg=gconjg)
    for key, (Ms, w) in _vertices(gold).items():
        want = 0j
        for x in range(Q.shape[2]):
            P = np.conj(A[:, :, x]) if gconjg else np.conj(A[:, :, x]).T
            for M in Ms:
                want += 1j * np.trace(P @ M @ Q[:, :, x] @ M) * w
        assert abs(got[key] - want) < 1e-12 * abs(want), key


def test_gconjg_variant_is_transposed_antiquark(oracle):
    """conj(A) = (A^T)^dag site by site: the _gconjg correlators of (Q, A)
    are the plain ones of (Q, A^T)."""
    rng = np.random.default_rng(43)
    Q, A = _random_props(oracle, rng, (3, 2, 2))
    At = np.asfortranarray(np.transpose(A, (1, 0, 2)))
    g = oracle.correlators(Q, A, gconjg=True)
    p = oracle.correlators(Q, At)
    for key in KEYS:
        assert abs(g[key] - p[key]) < 1e-13 * abs(p[key]), key
    # and they differ from the plain ones of (Q, A) for general propagators
    plain = oracle.correlators(Q, A)
    assert all(abs(g[k] - plain[k]) > 1e-3 * abs(plain[k]) for k in KEYS)


def test_colour_singlet_propagators(oracle):
    """Q = 1_3 (x) q(x), A = 1_3 (x) a(x) with 2x2 spin blocks: the colour
    trace of every octet vertex is sum_a Tr(T^a T^a) = 4 against Tr(1_3) = 3
    for the singlet, and the weights 6/8 and 6/24 against 1 and 1/3 make up
    the ratio, so O8(1S0) = O1(1S0) and O8(3S1) = O1(3S1)."""
    rng = np.random.default_rng(44)
    oracle.lattice_init((2, 2, 1), (0.05, 1.0, 1.0, 1.0))
    vol = 4
    Q = np.zeros((6, 6, vol), np.complex128, order="F")
    A = np.zeros_like(Q)
    for x in range(vol):
        q = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        a = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        Q[:, :, x] = np.kron(np.eye(3), q)
        A[:, :, x] = np.kron(np.eye(3), a)
    c = oracle.correlators(Q, A)
    assert abs(c["O8_1S0"] - c["O1_1S0"]) < 1e-13 * abs(c["O1_1S0"])
    assert abs(c["O8_3S1"] - c["O1_3S1"]) < 1e-13 * abs(c["O1_3S1"])
