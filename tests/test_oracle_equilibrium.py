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


def test_projection_sweep_in_the_adjoint_representation(oracle):
    """One projection sweep by a second route.  G is Hermitian traceless, so
    G = g^a T^a with g^a = 2 Re Tr(T^a G); U G U^+ has coordinates R(U) g with
    R(U)_ab = 2 Re Tr(T^a U T^b U^+); and 2 Re Tr(T^c A B) = d^abc alpha^a beta^b / 2
    for A = alpha^a T^a, B = beta^b T^b, with d^abc = 2 Tr({T^a, T^b} T^c).  Hence
    E^c_k(x) <- kappa a0 [d^abc (R(U_k(x)) g(x))_a g(x+k)_b / 2 - g^c(x)] + E^c_k(x),
    with x + k taken from the site coordinates.  A spacing a0 != 1 separates
    kappa from kappa a0; random links separate U G U^+ from U^+ G U."""
    ext, sp = (3, 4, 5), (0.07, 1.0, 0.9, 1.2)
    lat = _lat(oracle, ext, sp)
    U, E = oracle.box_init(3, 0.7, 1.0, 1.0)
    oracle.leapfrog(U, E, 1.0, 2)
    W = gf.links_sdrc(U)
    G = eo.gauss_matrices(lat, W, E)
    got = eo.project(lat, W, E, G)

    Tm = np.transpose(lat.T, (2, 0, 1))  # (8, 3, 3)
    g = 2 * np.real(np.einsum("arc,xcr->xa", Tm, G))
    assert np.allclose(G, np.einsum("xa,arc->xrc", g, Tm), atol=1e-13)  # premise: G in the algebra
    d = np.zeros((8, 8, 8))
    for a in range(8):
        for b in range(8):
            ac = Tm[a] @ Tm[b] + Tm[b] @ Tm[a]
            for c in range(8):
                d[a, b, c] = 2 * np.real(np.trace(ac @ Tm[c]))
    n0, n1, n2 = ext
    kdt = eo.KAPPA * sp[0]
    E = np.asarray(E)
    for s in range(lat.vol):
        x, y, z = s % n0, (s // n0) % n1, s // (n0 * n1)
        for k in range(3):
            q = [x, y, z]
            q[k] = (q[k] + 1) % ext[k]
            sk = q[0] + n0 * (q[1] + n1 * q[2])
            u = W[s, k]
            R = np.array([[2 * np.real(np.trace(Tm[a] @ u @ Tm[b] @ u.conj().T)) for b in range(8)] for a in range(8)])
            want = kdt * (0.5 * np.einsum("abc,a,b->c", d, R @ g[s], g[sk]) - g[s]) + E[:, k, s]
            assert np.allclose(got[:, k, s], want, rtol=0, atol=1e-13), (s, k)
