"""Index, neighbour and momentum maps: the oracle against brute-force
enumeration and closed forms, and the product's C-ABI lattice module against
the oracle (bit-exact; host-only calls, no GPU)."""
import itertools

import numpy as np
import pytest

LATTICES = [(1, 1, 1), (2, 2, 2), (3, 4, 5), (7, 1, 2), (16, 16, 16), (5, 6, 7)]


def _positions(ext):
    for z, y, x in itertools.product(range(1, ext[2] + 1), range(1, ext[1] + 1), range(1, ext[0] + 1)):
        yield (x, y, z)


@pytest.mark.parametrize("ext", LATTICES)
def test_index_bijection_and_neighbours(oracle, ext):
    oracle.lattice_init(ext, (0.1, 1.0, 1.0, 1.0))
    seen = set()
    for n, pos in enumerate(_positions(ext), start=1):
        idx = oracle.index(pos)
        # x fastest, 1-based (lattice.f90:883-898): enumeration order == index
        assert idx == n
        assert tuple(oracle.position(idx)) == pos
        seen.add(idx)
        for d in range(3):
            for sgn in (+1, -1):
                q = list(pos)
                q[d] = (q[d] - 1 + sgn) % ext[d] + 1
                assert oracle.neib(sgn * (d + 1), idx) == oracle.index(q)
        assert oracle.neib(0, idx) == idx
        neg = [(ext[d] - (pos[d] - 1)) % ext[d] + 1 for d in range(3)]
        assert oracle.negative_index(idx) == oracle.index(neg)
    assert seen == set(range(1, int(np.prod(ext)) + 1))


def test_momentum_closed_form(oracle):
    ext, sp = (5, 8, 3), (0.05, 1.0, 0.7, 1.3)
    oracle.lattice_init(ext, sp)
    for pos in _positions(ext):
        idx = oracle.index(pos)
        n = np.array(pos) - 1
        want = 2 * np.sin(np.pi * n / np.array(ext)) / np.array(sp[1:]) * np.exp(-1j * np.pi * n / np.array(ext))
        assert np.allclose(oracle.momentum(idx), want, rtol=1e-15, atol=1e-15)
        assert oracle.norm2_momentum(idx) == pytest.approx(np.linalg.norm(np.abs(want)), rel=1e-15, abs=1e-15)


@pytest.mark.parametrize("ext", [(3, 4, 5), (16, 16, 16), (1, 2, 9)])
def test_product_lattice_matches_oracle(oracle, nrq, lattice, ext):
    sp = (0.05, 1.0, 0.9, 1.1)
    oracle.lattice_init(ext, sp)
    lattice(ext, sp)
    assert nrq.lattice_size() == oracle.vol
    for idx in range(1, oracle.vol + 1):
        pos = oracle.position(idx)
        assert np.array_equal(nrq.GetLatticePosition(idx), pos)
        assert nrq.GetLatticeIndex(pos) == idx
        for d in (-3, -2, -1, 0, 1, 2, 3):
            assert nrq.GetNeib_G(d, idx) == oracle.neib(d, idx)
        assert nrq.GetNegativeLatticeIndex(idx) == oracle.negative_index(idx)
        assert np.allclose(nrq.GetMomentum_G(idx), oracle.momentum(idx), rtol=1e-15, atol=1e-15)


def test_product_lattice_errors(nrq, lattice):
    with pytest.raises(nrq.NRQError) as e:
        nrq.GetLatticePosition(1)
    assert e.value.code == -2  # not initialised
    lattice((4, 4, 4), (0.05, 1, 1, 1))
    with pytest.raises(nrq.NRQError) as e:
        nrq.lattice_init((4, 4, 4), (0.05, 1, 1, 1))
    assert e.value.code == -2  # already initialised
    for bad in (0, 65, -3):
        with pytest.raises(nrq.NRQError):
            nrq.GetLatticePosition(bad)
    with pytest.raises(nrq.NRQError):
        nrq.GetNeib_G(4, 1)


def test_product_lattice_rejects_bad_input(nrq):
    with pytest.raises(nrq.NRQError):
        nrq.lattice_init((0, 4, 4), (0.05, 1, 1, 1))
    with pytest.raises(nrq.NRQError):
        nrq.lattice_init((4, 4, 4), (0.05, -1, 1, 1))
