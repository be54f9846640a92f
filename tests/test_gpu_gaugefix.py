"""GPU parity of Coulomb gauge fixing and the transverse AA/EE correlators
(nrq_gaugefix.hip) against the CPU restatement oracle/gaugefix_oracle.py.

Bars: gauge fixing is an iteration whose FP64 trajectories differ only by FFT
rounding, so the transformation counts agree to +-1 and the links to 1e-8;
the correlators of the same links agree to 1e-10 relative to their maximum.
"""
import numpy as np
import pytest

from oracle import gaugefix_oracle as gf

pytestmark = pytest.mark.gpu

SP = (0.05, 1.0, 1.0, 1.0)


def _say(msg):
    print(msg, flush=True)


def _setup(nrq, oracle, lattice, ext, sp=SP):
    lattice(ext, sp)
    oracle.lattice_init(ext, sp)
    return gf.Lattice(ext, sp, oracle.generators())


def _evolved(oracle, seed=9, steps=5):
    U, E = oracle.box_init(seed, 0.5, 1.0, 1.0)
    oracle.leapfrog(U, E, 1.0, steps)
    return U, E


def test_gaugefix_matches_oracle(nrq, oracle, lattice):
    ext = (8, 6, 5)
    _say("Coulomb gauge fixing 8x6x5 FP64")
    lat = _setup(nrq, oracle, lattice, ext)
    U, E = _evolved(oracle)
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    its, dev = g.CoulombGaugefixing(1e-10, 0.08, 600)
    V, done, odev = gf.coulomb_gauge_fix(lat, U, 1e-10, 0.08, 600)
    _say(f"transformations: gpu {its}, oracle {done}; deviation gpu {dev:.3e}, oracle {odev:.3e}")
    assert 0 < its < 600 and dev <= 1e-10
    assert abs(its - done) <= 1
    Ug, Eg = g.download()
    assert np.max(np.abs(Ug - V)) < 1e-8
    assert np.array_equal(Eg, E)  # E is not transformed, as in the reference
    # gauge invariants
    assert np.allclose(gf.plaquettes(lat, Ug), gf.plaquettes(lat, U), atol=1e-11)
    assert g.GetEnergy() == pytest.approx(oracle.energy(U, E), rel=1e-12)


def test_gaugefix_iteration_cap_and_defaults(nrq, oracle, lattice):
    ext = (6, 4, 4)
    _say("gauge fixing cap and defaults")
    lat = _setup(nrq, oracle, lattice, ext)
    U, E = _evolved(oracle, seed=3)
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    its, dev = g.CoulombGaugefixing(1e-14, 0.08, 3)
    assert its == 3 and dev > 1e-14
    V, done, odev = gf.coulomb_gauge_fix(lat, U, 1e-14, 0.08, 3)
    assert done == 3 and dev == pytest.approx(odev, rel=1e-8)
    assert np.max(np.abs(g.download()[0] - V)) < 1e-12
    h = nrq.GaugeConfiguration()
    h.upload(U, E)
    its, dev = h.CoulombGaugefixing()  # the reference's defaults: 1e-8, 0.08, 100
    assert dev <= 1e-8 or its == 100
    with pytest.raises(nrq.NRQError):
        h.CoulombGaugefixing(-1.0)


def test_pure_gauge_fixed_to_unity_gpu(nrq, oracle, lattice):
    ext = (8, 8, 6)
    _say("pure gauge 8x8x6")
    lat = _setup(nrq, oracle, lattice, ext)
    rng = np.random.default_rng(21)
    h = np.stack([oracle.group_exp(rng.normal(size=8) * 0.3) for _ in range(lat.vol)])
    W = np.zeros((lat.vol, 3, 3, 3), np.complex128)
    for i in range(3):
        W[:, i] = h @ gf.dag(h[lat.up[i]])
    U = gf.links_fortran(W)
    _, E = oracle.cold()
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    its, dev = g.CoulombGaugefixing(1e-11, 0.08, 1000)
    assert its > 0 and dev <= 1e-11
    assert np.max(np.abs(gf.links_sdrc(g.download()[0]) - np.eye(3))) < 1e-8


def test_gaugefix_fp32_storage(nrq, oracle, lattice):
    ext = (8, 6, 5)
    _say("gauge fixing FP32 storage")
    lat = _setup(nrq, oracle, lattice, ext)
    U, E = _evolved(oracle)
    g = nrq.GaugeConfiguration(nrq.NRQ_PRECISION_FP32)
    g.upload(U, E)
    its, dev = g.CoulombGaugefixing(1e-6, 0.08, 600)
    assert dev <= 1e-6
    V, _, _ = gf.coulomb_gauge_fix(lat, U, 1e-6, 0.08, 600)
    Ug = g.download()[0]
    assert np.max(np.abs(Ug - V)) < 1e-4
    assert np.allclose(gf.plaquettes(lat, Ug), gf.plaquettes(lat, U), atol=1e-4)


def test_correlators_match_oracle(nrq, oracle, lattice):
    ext = (8, 6, 5)
    sp = (0.05, 1.0, 0.9, 1.1)
    _say("transverse correlators 8x6x5, anisotropic")
    lat = _setup(nrq, oracle, lattice, ext, sp)
    U, E = _evolved(oracle)
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    g.CoulombGaugefixing(1e-9, 0.08, 500)
    Ug, Eg = g.download()
    aa, ee = g.transverse_correlators()
    raa, ree = gf.transverse_correlators(lat, Ug, Eg)
    assert np.max(np.abs(aa - raa)) < 1e-10 * np.max(np.abs(raa))
    assert np.max(np.abs(ee - ree)) < 1e-10 * np.max(np.abs(ree))
    assert np.array_equal(g.GetTransverseAACorrelator(), aa)
    assert np.array_equal(g.GetTransverseEECorrelator(), ee)
    # gluon distribution as the driver writes it: |p| and sqrt(AA EE)
    p = nrq.momentum_norms()
    assert np.allclose(p, lat.pnorm, atol=1e-13)
    assert nrq.GetMaxNorm2Momentum() == pytest.approx(lat.pmax, rel=1e-15)
    f = np.sqrt(aa * ee)
    assert np.all(np.isfinite(f)) and f[0] == 0.0


def test_group_log_far_from_identity(nrq, oracle, lattice):
    """On a 2x2x2 lattice every link enters AA with visible weight: links with
    eigenphases up to +-pi (random algebra vectors of norm up to ~5) exercise
    the eigen-decomposition, including its second Hermitian pencil."""
    ext = (2, 2, 2)
    _say("group log far from identity 2x2x2")
    lat = _setup(nrq, oracle, lattice, ext)
    rng = np.random.default_rng(31)
    for trial in range(6):
        U, E = oracle.cold()
        for s in range(lat.vol):
            for i in range(3):
                U[:, :, i, s] = oracle.group_exp(rng.normal(size=8) * (0.5 + trial * 0.5))
        E[:] = rng.normal(size=E.shape)
        g = nrq.GaugeConfiguration()
        g.upload(U, E)
        aa, ee = g.transverse_correlators()
        raa, ree = gf.transverse_correlators(lat, U, E)
        assert np.max(np.abs(aa - raa)) < 1e-10 * np.max(np.abs(raa)), trial
        assert np.max(np.abs(ee - ree)) < 1e-12 * np.max(np.abs(ree)), trial
