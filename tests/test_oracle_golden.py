"""Pin the oracle (and the product's host RNG) against the reference's own
compiled arithmetic: oracle/_ref when it is built, otherwise the committed
fixtures tests/golden/ref_vectors.npz written from it by
tests/golden/make_golden.py."""
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden" / "ref_vectors.npz"


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.mark.parametrize("level", [1, 2])
@pytest.mark.parametrize("seed", [1, 2, 3, 100, 131, 2147483647])
def test_ranlxd_stream_bit_exact(oracle, gold, level, seed):
    oracle.ranlxd_init(level, seed)
    r = oracle.ranlxd(3000)
    assert np.array_equal(r, gold[f"ranlxd_l{level}_s{seed}"])
    assert np.array_equal(oracle.ranlxd_get(), gold[f"ranlxd_state_l{level}_s{seed}"])


@pytest.mark.parametrize("seed", [1, 2, 3, 100, 131])
def test_product_rng_stream_bit_exact(gold, seed):
    """The library's own host ranlxd (random::InitModule -> level 2)."""
    import nrqcd_amd as nrq

    nrq.random_init(seed)
    assert np.array_equal(nrq.random_uniform(3000), gold[f"ranlxd_l2_s{seed}"])


@pytest.mark.parametrize("seed", [2, 987654321])
def test_product_rng_long_stream_in_uneven_draws(oracle, seed):
    """200,000 numbers (about 6,600 refreshes of the 12-number state) drawn in
    chunks of uneven length equal the oracle's stream, itself pinned by the
    golden vectors above; the library keeps the state as 48-bit integer
    numerators, so this checks that form over many carries and wraps."""
    import nrqcd_amd as nrq

    nrq.random_init(seed)
    oracle.ranlxd_init(2, seed)
    sizes = [1, 11, 12, 13, 5, 24, 25, 7, 100, 997]
    got, n = [], 0
    k = 0
    while n < 200000:
        m = min(sizes[k % len(sizes)], 200000 - n)
        got.append(nrq.random_uniform(m))
        n += m
        k += 1
    assert np.array_equal(np.concatenate(got), oracle.ranlxd(200000))


def test_product_rng_rejects_bad_seed():
    import nrqcd_amd as nrq

    with pytest.raises(nrq.NRQError):
        nrq.random_init(0)
    with pytest.raises(nrq.NRQError):
        nrq.random_init(2 ** 31)


def test_box_muller_from_reference_stream(oracle, gold):
    """random.f90:238-253: sqrt(-2 ln r1) (cos 2 pi r2 + i sin 2 pi r2) on
    consecutive pairs of the ranlxd stream."""
    oracle.ranlxd_init(2, 1)
    z = oracle.normal_cmplx(1000)
    r = gold["ranlxd_l2_s1"][:2000]
    want = np.sqrt(-2 * np.log(r[0::2])) * np.exp(2j * np.pi * r[1::2])
    assert np.allclose(z, want, rtol=1e-14, atol=1e-15)


def test_generators_match_reference(oracle, gold):
    assert np.array_equal(oracle.generators(), gold["su3_generators"])


def test_pauli_matches_reference(oracle, gold):
    for k in range(1, 4):
        assert np.array_equal(oracle.pauli(k), gold["su2_generators"][:, :, k - 1])


def test_group_exp_matches_reference(oracle, gold):
    for a, u in zip(gold["exp_alpha"], gold["exp_U"]):
        got = oracle.group_exp(a)
        assert np.max(np.abs(got - u)) < 2e-14 * max(1.0, np.abs(a).sum()), a


def test_algebra_maps_match_reference(oracle, gold):
    for m, coord in zip(gold["alg_M"], gold["alg_coord"]):
        got = [oracle.algebra_coordinate(a, m) for a in range(1, 9)]
        assert np.allclose(got, coord, rtol=0, atol=1e-14)
    for a, m in zip(gold["exp_alpha"][:16], gold["alg_matrix"]):
        assert np.allclose(oracle.algebra_matrix(a), m, rtol=0, atol=1e-15 * max(1, np.abs(a).max()))


def test_kron_matches_reference(oracle, gold):
    assert np.array_equal(oracle.kron(gold["kron_A"], gold["kron_B"]), gold["kron_AB"])


def test_live_reference_modules(oracle, refmods):
    """Direct comparison with the compiled reference modules where present."""
    rng = np.random.default_rng(7)
    for seed in (5, 77, 12345):
        refmods.ranlxd_init(2, seed)
        oracle.ranlxd_init(2, seed)
        assert np.array_equal(refmods.ranlxd(10007), oracle.ranlxd(10007))
    for _ in range(50):
        a = rng.normal(size=8) * rng.choice([1e-4, 0.1, 1.0, 4.0])
        assert np.max(np.abs(refmods.group_exp(a) - oracle.group_exp(a))) < 2e-14 * max(1, np.abs(a).sum())
        m = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        for g in range(1, 9):
            assert abs(refmods.algebra_coordinate(g, m) - oracle.algebra_coordinate(g, m)) < 1e-14
