"""Pinning the CPU restatement of EquilibriumInit (oracle/equilibrium_oracle.py,
src/gaugeconfiguration_su3.f90:380-814); no GPU."""
import numpy as np
import pytest

from oracle import equilibrium_oracle as eo
from oracle import gaugefix_oracle as gf


def _lat(oracle, ext, sp=(0.05, 1.0, 1.0, 1.0)):
    oracle.lattice_init(ext, sp)
    return gf.Lattice(ext, sp, oracle.generators())


@pytest.mark.parametrize("ext,sp", [((4, 4, 4), (0.05, 1.0, 1.0, 1.0)), ((3, 5, 4), (0.1, 1.0, 0.8, 1.2))])
def test_gauss_deviation_is_the_c_oracle_measure(oracle, ext, sp):
    lat = _lat(oracle, ext, sp)
    U, E = oracle.box_init(3, 0.7, 1.0, 1.0)
    oracle.leapfrog(U, E, 1.0, 3)
    G = eo.gauss_matrices(lat, gf.links_sdrc(U), E)
    assert eo.deviation(lat, G) == pytest.approx(oracle.gauss_max(U, E), rel=1e-12)


def test_gauss_law_configuration_is_a_fixed_point(oracle):
    """Cold links with the box IC's transverse E satisfy the Gauss law, so the
    projection returns E unchanged."""
    lat = _lat(oracle, (4, 5, 3))
    _, E = oracle.box_init(4, 0.8, 1.0, 1.0)
    U, _ = oracle.cold()
    W = gf.links_sdrc(U)
    G = eo.gauss_matrices(lat, W, E)
    assert eo.deviation(lat, G) < 1e-12
    assert np.max(np.abs(eo.project(lat, W, E, G) - E)) < 1e-13


def test_charge_density_enters_the_gauss_law(oracle):
    """E = 0 and a static colour charge rho: G = -sum_a T^a Tr(T^a C) / a0 with
    C = diag(rho) / (a1 a2 a3), the traceless part of -C / 2 / a0."""
    sp = (0.05, 1.0, 0.9, 1.1)
    lat = _lat(oracle, (3, 3, 3), sp)
    U, E = oracle.cold()
    rng = np.random.default_rng(2)
    rho = rng.normal(size=(3, lat.vol))
    G = eo.gauss_matrices(lat, gf.links_sdrc(U), E, rho)
    for x in range(lat.vol):
        C = np.diag(rho[:, x]) / (sp[1] * sp[2] * sp[3])
        want = -(C - np.trace(C) / 3 * np.eye(3)) / 2 / sp[0]
        assert np.allclose(G[x], want, atol=1e-13)


def test_electric_field_draws(oracle):
    """The first draw of an outer iteration: 16 uniforms per site and
    direction, Box-Muller real parts times sigma a0, before any projection."""
    lat = _lat(oracle, (3, 3, 3))
    oracle.ranlxd_init(2, 9)
    beta = 12.0
    U, E, rec = eo.equilibrium_init(oracle, lat, beta, 1, 0, tol=1e300)
    assert rec[0][0] == 0  # no projection with this tolerance
    oracle.ranlxd_init(2, 9)
    u = oracle.ranlxd(48 * lat.vol).reshape(lat.vol, 3, 16)
    want = np.sqrt(4 / beta) * np.sqrt(-2 * np.log(u[:, :, 0::2])) * np.cos(2 * np.pi * u[:, :, 1::2]) * lat.a[0]
    assert np.allclose(E, np.transpose(want, (2, 1, 0)), rtol=1e-15, atol=0)


def test_projection_reaches_the_tolerance(oracle):
    lat = _lat(oracle, (4, 4, 4))
    oracle.ranlxd_init(2, 5)
    U, E, rec = eo.equilibrium_init(oracle, lat, 16.0, 1, 3)
    assert rec[0][0] > 0
    assert rec[0][2] <= 2 * eo.DEFAULT_TOL  # the leapfrog keeps the Gauss law
    assert eo.DEFAULT_TOL == 0.0010000000474974513
