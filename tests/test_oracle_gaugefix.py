"""Pinning the CPU restatement of Coulomb gauge fixing and the transverse
correlators (oracle/gaugefix_oracle.py); no GPU."""
import numpy as np
import pytest

from oracle import gaugefix_oracle as gf
from oracle.pyoracle import Oracle, RefModules

SP = (0.05, 1.0, 1.0, 1.0)


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def _lat(orc, ext, sp=SP):
    orc.lattice_init(ext, sp)
    return gf.Lattice(ext, sp, orc.generators())


def _random_su3(orc, rng, scale):
    return orc.group_exp(rng.normal(size=8) * scale)


@pytest.mark.skipif(not RefModules.available(), reason="oracle/_ref not built")
def test_group_log_matches_compiled_reference(orc):
    """GetGroupLog of the reference (compiled from its own su3/matrixoperations
    sources) on near-identity links, links with eigenphases near +-pi and
    products of random group elements."""
    ref = RefModules()
    lat = _lat(orc, (2, 2, 2))
    rng = np.random.default_rng(11)
    mats = [_random_su3(orc, rng, s) for s in (0.01, 0.3, 1.0, 2.0, 3.0) for _ in range(6)]
    mats += [_random_su3(orc, rng, 2.0) @ _random_su3(orc, rng, 2.0) for _ in range(6)]
    M = np.stack(mats)
    ours = gf.group_log_coords(lat, M)
    for k, m in enumerate(mats):
        assert np.allclose(ours[k], ref.group_log(m), rtol=0, atol=1e-11), k


def test_group_log_inverts_exp(orc):
    """For |theta| < pi the principal log returns the algebra coordinates the
    exponential was built from (mathematics, independent of the reference)."""
    lat = _lat(orc, (2, 2, 2))
    rng = np.random.default_rng(12)
    alphas = rng.normal(size=(20, 8)) * 0.4
    M = np.stack([orc.group_exp(a) for a in alphas])
    assert np.allclose(gf.group_log_coords(lat, M), alphas, atol=1e-12)


@pytest.mark.skipif(not RefModules.available(), reason="oracle/_ref not built")
def test_expah_matches_compiled_reference_off_traceless():
    """The divergence of the gauge-fixing step is skew-Hermitian but not
    traceless; the exponential agrees with the reference's ExpAH there."""
    ref = RefModules()
    rng = np.random.default_rng(13)
    for _ in range(10):
        X = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        A = (X - X.conj().T) / 2 * 0.7
        assert abs(np.trace(A)) > 1e-3
        assert np.allclose(gf.expah(A[None])[0], ref.expah(A), atol=1e-13)


def test_x2p_p2x_match_c_oracle(orc):
    ext = (4, 3, 5)
    lat = _lat(orc, ext, (0.05, 1.0, 0.8, 1.3))
    rng = np.random.default_rng(14)
    f = rng.normal(size=lat.vol) + 1j * rng.normal(size=lat.vol)
    assert np.allclose(lat.x2p(f), orc.x2p(f), atol=1e-12)
    assert np.allclose(lat.p2x(f), orc.p2x(f), atol=1e-12)
    assert np.allclose(lat.p2x(lat.x2p(f)), f, atol=1e-12)


def test_momenta_match_c_oracle(orc):
    ext = (4, 3, 5)
    lat = _lat(orc, ext, (0.05, 1.0, 0.8, 1.3))
    for idx in range(1, lat.vol + 1, 7):
        assert np.allclose(lat.p[idx - 1], orc.momentum(idx), atol=1e-14)
        assert abs(lat.pnorm[idx - 1] - orc.norm2_momentum(idx)) < 1e-14


def test_pure_gauge_is_fixed_to_unity(orc):
    """U_i(x) = h(x) h(x+i)^+ is a gauge transform of U = 1; Coulomb gauge
    fixing must bring it back to (a constant rotation of) the identity."""
    ext = (4, 4, 4)
    lat = _lat(orc, ext)
    rng = np.random.default_rng(15)
    h = np.stack([_random_su3(orc, rng, 0.3) for _ in range(lat.vol)])
    W = np.zeros((lat.vol, 3, 3, 3), np.complex128)
    for i in range(3):
        W[:, i] = h @ gf.dag(h[lat.up[i]])
    U = gf.links_fortran(W)
    V, done, dev = gf.coulomb_gauge_fix(lat, U, tol=1e-10, maxit=500)
    assert done > 0 and dev <= 1e-10
    assert np.max(np.abs(gf.links_sdrc(V) - np.eye(3))) < 1e-8


def test_gauge_fixing_keeps_plaquettes_and_unitarity(orc):
    ext = (4, 4, 6)
    lat = _lat(orc, ext)
    U, E = orc.box_init(5, 0.5, 1.0, 1.0)
    V, done, dev = gf.coulomb_gauge_fix(lat, U, tol=1e-9, maxit=300)
    assert done > 0 and dev <= 1e-9
    assert np.allclose(gf.plaquettes(lat, V), gf.plaquettes(lat, U), atol=1e-11)
    W = gf.links_sdrc(V)
    assert np.max(np.abs(W @ gf.dag(W) - np.eye(3))) < 1e-12
    # the fixed configuration satisfies the lattice Coulomb condition
    assert gf.deviation(gf.divergence(lat, W)).max() <= 1e-9


def test_transverse_plane_wave_closed_form(orc):
    """E^1_y(x) = e0 cos(2 pi x / n0): only the momenta +-(1,0,0) carry it, with
    |x2p|^2 = (a1 a2 a3 V / 2)^2 and P_yy = 1/2, so
    EE(+-p) = (a1 a2 a3 V / 2)^2 / 2 / (V a1 a2 a3) / 8 * (e0 / (a2 a0))^2.
    A longitudinal wave (E along x) gives zero."""
    ext = (6, 4, 4)
    sp = (0.05, 1.0, 1.0, 1.0)
    lat = _lat(orc, ext, sp)
    U, E = orc.cold()
    n0 = ext[0]
    x = np.arange(lat.vol) % n0
    e0 = 0.3
    E[0, 1, :] = e0 * np.cos(2 * np.pi * x / n0)
    aa, ee = gf.transverse_correlators(lat, U, E)
    fwd = sp[1] * sp[2] * sp[3]
    want = (fwd * lat.vol / 2) ** 2 / 2 / lat.vol_phys / 8 * (e0 / (sp[2] * sp[0])) ** 2
    hit = [1, n0 - 1]  # momentum indices (1,0,0) and (n0-1,0,0)
    assert np.allclose(ee[hit], want, rtol=1e-12)
    rest = np.delete(ee, hit)
    assert np.max(np.abs(rest)) < 1e-12 * want
    assert np.max(np.abs(aa)) < 1e-20  # cold links: A = 0
    E[:] = 0
    E[0, 0, :] = e0 * np.cos(2 * np.pi * x / n0)
    _, ee_l = gf.transverse_correlators(lat, U, E)
    assert np.max(np.abs(ee_l)) < 1e-12 * want


def test_aa_of_small_plane_wave_links(orc):
    """Links U_y(x) = exp(i a_y A cos(2 pi x / n0) T^1) have group-log
    coordinates a_y A cos(...), so AA follows the same closed form as EE."""
    ext = (6, 4, 4)
    sp = (0.05, 1.0, 1.0, 1.0)
    lat = _lat(orc, ext, sp)
    U, E = orc.cold()
    n0 = ext[0]
    amp = 0.2
    for s in range(lat.vol):
        alpha = np.zeros(8)
        alpha[0] = sp[2] * amp * np.cos(2 * np.pi * (s % n0) / n0)
        U[:, :, 1, s] = orc.group_exp(alpha)
    aa, _ = gf.transverse_correlators(lat, U, E)
    fwd = sp[1] * sp[2] * sp[3]
    want = (fwd * lat.vol / 2) ** 2 / 2 / lat.vol_phys / 8 * amp ** 2
    assert np.allclose(aa[[1, n0 - 1]], want, rtol=1e-10)


def test_projector_is_real_part_off_axis(orc):
    """The reference's projector is declared real, so only
    Re[(delta_ij - conj(p_i) p_j / |p|^2) / 2] enters.  For a wave along
    (1, 1, 0) on a lattice with n0 != n1 the components of the
    backward-derivative momentum have different phases.  E^1_x = cos(k.x),
    E^1_y = sin(k.x) give F_y = -+i F_x at +-k, |F|^2 = (a1 a2 a3 V / 2 / a0)^2,
    and with the real projector sum_ij F_i P_ij conj(F_j) = (P_xx + P_yy) |F|^2
    = |F|^2 / 2 at both momenta; the complex projector would add
    +- |F|^2 Im(conj(p_x) p_y) / |p|^2."""
    ext = (6, 4, 3)
    sp = (0.05, 1.0, 1.0, 1.0)
    lat = _lat(orc, ext, sp)
    U, E = orc.cold()
    s = np.arange(lat.vol)
    x, y = s % ext[0], (s // ext[0]) % ext[1]
    ph = 2 * np.pi * (x / ext[0] + y / ext[1])
    E[0, 0, :] = np.cos(ph)
    E[0, 1, :] = np.sin(ph)
    _, ee = gf.transverse_correlators(lat, U, E)
    F2 = (sp[1] * sp[2] * sp[3] * lat.vol / 2 / sp[0]) ** 2
    want = F2 / 2 / lat.vol_phys / 8
    for k in (1 + ext[0] * 1, (ext[0] - 1) + ext[0] * (ext[1] - 1)):  # (1,1,0) and (-1,-1,0)
        p = orc.momentum(k + 1)
        assert abs(np.imag(np.conj(p[0]) * p[1])) / np.sum(np.abs(p) ** 2) > 1e-2
        assert ee[k] == pytest.approx(want, rel=1e-12)
