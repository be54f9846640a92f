"""GPU parity of EquilibriumInit (nrq_equilibrium.hip) against
oracle/equilibrium_oracle.py: the same number of projection sweeps, links
and E within 1e-10 after the projections and the leapfrog steps, energies
and Gauss deviations within 1e-10 relative."""
import numpy as np
import pytest

from oracle import equilibrium_oracle as eo
from oracle import gaugefix_oracle as gf

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("ext,sp,beta,nef,neq,tol,charge", [
    ((4, 4, 4), (0.05, 1.0, 1.0, 1.0), 16.0, 2, 5, None, False),
    # an even and an odd number of leapfrog steps leave the links in either
    # of the gauge field's two buffers before the second pass
    ((4, 4, 4), (0.05, 1.0, 1.0, 1.0), 16.0, 2, 4, None, False),
    ((3, 4, 5), (0.1, 1.0, 0.8, 1.2), 8.0, 1, 3, 5e-3, True),
])
def test_equilibrium_init_matches_oracle(nrq, oracle, lattice, ext, sp, beta, nef, neq, tol, charge):
    _say(f"EquilibriumInit {ext} beta {beta} x{nef}, {neq} steps, tol {tol}, charge {charge}")
    lattice(ext, sp)
    oracle.lattice_init(ext, sp)
    lat = gf.Lattice(ext, sp, oracle.generators())
    rho = None
    if charge:
        rho = np.random.default_rng(3).normal(size=(3, lat.vol)) * 0.2
        rho -= rho.mean(axis=1, keepdims=True)
    nrq.random_init(5)
    oracle.ranlxd_init(2, 5)
    g = nrq.GaugeConfiguration()
    en, dev, nproj = g.EquilibriumInit(beta, nef, neq, tol, rho)
    U, E, rec = eo.equilibrium_init(oracle, lat, beta, nef, neq, tol, rho)
    assert list(nproj) == [r[0] for r in rec] and min(nproj) > 0
    Ug, Eg = g.download()
    assert np.max(np.abs(Ug - U)) < 1e-10
    assert np.max(np.abs(Eg - E)) < 1e-10 * max(1.0, np.max(np.abs(E)))
    assert np.allclose(en, [r[1] for r in rec], rtol=1e-10)
    assert np.allclose(dev, [r[2] for r in rec], rtol=1e-8)
    # the same random stream position afterwards
    assert np.array_equal(nrq.random_uniform(5), oracle.ranlxd(5))


def test_equilibrium_init_rejects_bad_arguments(nrq, lattice):
    lattice((4, 4, 4), (0.05, 1.0, 1.0, 1.0))
    nrq.random_init(1)
    g = nrq.GaugeConfiguration()
    with pytest.raises(nrq.NRQError):
        g.EquilibriumInit(-1.0, 1, 1)
    with pytest.raises(nrq.NRQError):
        g.EquilibriumInit(16.0, -1, 1)
