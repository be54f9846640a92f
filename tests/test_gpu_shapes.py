"""Every launch shape of the leapfrog and of the CN stencil kernels against
the oracle, and the launch caches against lattice changes.

The autotuner picks one leapfrog shape by timing; any shape it may pick must
give the oracle's answer.  Each shape is pinned with nrq_leapfrog_set_shape
and run from the oracle's state: FP64 links and E within 1e-12 relative after
10 steps (the two-row shapes rebuild the third row of neighbour links, exact
for SU(3) up to rounding), FP32 storage within 2e-5.  Each CN shape
(fields per thread, tiling, XCD remap) runs one CN step that must match the
oracle's within 1e-10 (both solve to true relative residual 1e-12).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SP = (0.05, 1.0, 1.0, 1.0)
C1 = [1.0, 0.0, 0.0, 0.0]


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture()
def pinned(nrq):
    """Restores the autotuner's choice after the test."""
    yield
    nrq.leapfrog_set_shape(-1)


@pytest.mark.parametrize("ext", [(5, 6, 7), (16, 16, 16), (64, 64, 64)])
def test_every_leapfrog_shape_matches_oracle_fp64(nrq, oracle, lattice, pinned, ext):
    lattice(ext, SP)
    oracle.lattice_init(ext, SP)
    U0, E0 = oracle.box_init(2, 0.5, 1.0, 1.0)
    U, E = U0.copy(order="F"), E0.copy(order="F")
    oracle.leapfrog(U, E, 1.0, 10)
    shapes = nrq.leapfrog_shapes()
    tried = 0
    g = nrq.GaugeConfiguration()
    for k, sh in enumerate(shapes):
        if not sh["applies"]:
            continue
        _say(f"leapfrog shape {k} {sh} on {ext} fp64")
        nrq.leapfrog_set_shape(k)
        g.upload(U0, E0)
        g.Update(1.0, 10)
        Ug, Eg = g.download()
        assert _rel(Ug, U) < 1e-12, k
        assert _rel(Eg, E) < 1e-12, k
        tried += 1
    assert tried >= 9


@pytest.mark.parametrize("ext,sp", [((16, 16, 16), SP), ((32, 16, 24), (0.1, 1.0, 0.8, 1.3))])
def test_every_leapfrog_shape_matches_oracle_fp32(nrq, oracle, lattice, pinned, ext, sp):
    lattice(ext, sp)
    oracle.lattice_init(ext, sp)
    U0, E0 = oracle.box_init(3, 0.5, 1.0, 1.0)
    U, E = U0.copy(order="F"), E0.copy(order="F")
    oracle.leapfrog(U, E, 1.0, 10)
    g = nrq.GaugeConfiguration(nrq.NRQ_PRECISION_FP32)
    tried = []
    for k, sh in enumerate(nrq.leapfrog_shapes()):
        if not sh["applies_fp32"]:
            continue
        tried.append(k)
        _say(f"leapfrog shape {k} on {ext} {sp} fp32")
        nrq.leapfrog_set_shape(k)
        g.upload(U0, E0)
        g.Update(1.0, 10)
        Ug, Eg = g.download()
        assert np.max(np.abs(Ug - U)) < 2e-5, k
        assert np.max(np.abs(Eg - E)) < 2e-5 * max(1.0, np.max(np.abs(E))), k
    # the LDS-tiled shapes are among them
    assert any(nrq.leapfrog_shapes()[k]["tiling"].startswith("LDS") for k in tried)


def test_pinned_shape_that_does_not_apply_is_an_error(nrq, lattice, pinned):
    ext = (5, 6, 7)  # the 16-wide tiles need an x extent of at least 16
    _say("pinned shape that does not apply")
    lattice(ext, SP)
    shapes = nrq.leapfrog_shapes()
    k = next(i for i, sh in enumerate(shapes) if not sh["applies"])
    nrq.leapfrog_set_shape(k)
    assert nrq.leapfrog_get_shape() == k
    g = nrq.GaugeConfiguration()
    with pytest.raises(nrq.NRQError):
        g.Update(1.0, 2)
    with pytest.raises(nrq.NRQError):
        nrq.leapfrog_set_shape(len(shapes))


def test_autotuned_shape_matches_oracle(nrq, oracle, lattice, pinned):
    """The tuner's own pick (validated against the default shape) gives the
    oracle's answer at an anisotropic spacing, energy included."""
    ext = (32, 16, 16)
    sp = (0.1, 1.0, 0.8, 1.3)
    _say("autotuned shape, anisotropic 32x16x16")
    nrq.leapfrog_set_shape(-1)
    lattice(ext, sp)
    oracle.lattice_init(ext, sp)
    U, E = oracle.box_init(5, 0.6, 1.0, 1.0)
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    g.Update(1.0, 10)
    oracle.leapfrog(U, E, 1.0, 10)
    Ug, Eg = g.download()
    assert _rel(Ug, U) < 1e-12 and _rel(Eg, E) < 1e-12
    assert g.GetEnergy() == pytest.approx(oracle.energy(U, E), rel=1e-12)
    assert g.GetGaussMax() == pytest.approx(oracle.gauss_max(U, E), rel=1e-11)
    _say(f"tuned shape: {g.launch_shape()}")


CN_SHAPES = ([(nc, 4, t, x, nrq_kind) for nc in (1, 2, 4) for t in (0, 1, 2, 3, 4) for x in (1, 8)
              for nrq_kind in ("GMRES_FP64",)]
             + [(2, nc, t, 8, "CHEBYSHEV_MIXED") for nc in (1, 2, 4) for t in (0, 1, 2, 3, 4)])


def test_every_cn_shape_matches_oracle(nrq, oracle, lattice):
    ext = (16, 6, 5)
    sp = (0.05, 1.0, 0.9, 1.1)
    lattice(ext, sp)
    oracle.lattice_init(ext, sp)
    U, E = oracle.box_init(8, 0.5, 1.0, 1.0)
    oracle.leapfrog(U, E, 1.0, 5)
    Q, A = oracle.point_prop(1), oracle.point_prop(33)
    oracle.cn_update(U, Q, A, 2.0, C1, 1.0, 1e-13)
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    default = nrq.cn_get_shape()
    nrq.nrqcd_init(2.0, C1, 1.0, 1e-12)
    try:
        for nc64, nc32, tiling, xcd, kind in CN_SHAPES:
            _say(f"CN shape nc64={nc64} nc32={nc32} tiling={tiling} xcd={xcd} solver {kind}")
            nrq.cn_set_shape(nc64, nc32, tiling, xcd)
            nrq.nrqcd_set_solver(getattr(nrq, "NRQ_SOLVER_" + kind), 30, 10000)
            h = nrq.NRQCDField()
            h.InitSinglePoint(1, 33)
            st = h.Update(g, 2.0, C1)
            assert st.max_true_relres <= 1e-12
            Qg, Ag = h.download()
            assert _rel(Qg, Q) < 1e-10 and _rel(Ag, A) < 1e-10
        with pytest.raises(nrq.NRQError):
            nrq.cn_set_shape(3, 4, 0, 8)
    finally:
        nrq.cn_set_shape(*default)
        nrq.nrqcd_finalize()


def test_batch_graph_follows_a_change_of_spacings(nrq, lattice, monkeypatch):
    """The hipGraph of a batched leapfrog carries the lattice by value.  After
    the lattice module is re-initialised with the same extents and other
    spacings (new configurations, possibly at the same addresses), the batch
    must use the new spacings: it equals separate updates bit for bit."""
    monkeypatch.setenv("NRQ_GRAPH", "1")
    ext = (8, 6, 6)
    for sp in ((0.05, 1.0, 1.0, 1.0), (0.1, 1.0, 0.7, 1.4)):
        _say(f"batch graph at spacings {sp}")
        lattice(ext, sp)
        ens, ref = [], []
        for k in range(3):
            nrq.random_init(60 + k)
            g = nrq.GaugeConfiguration()
            g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
            ens.append(g)
            ref.append(g.copy())
        for _ in range(2):  # capture, then replay
            nrq.leapfrog_batch(ens, 1.0, 4)
            for r in ref:
                r.Update(1.0, 4)
        for a, b in zip(ens, ref):
            Ua, Ea = a.download()
            Ub, Eb = b.download()
            assert np.array_equal(Ua, Ub) and np.array_equal(Ea, Eb), sp
        for x in ens + ref:
            x.Deallocate()


def test_fused_launch_timer_counts_the_fused_passes(nrq, lattice):
    """The kernel timer behind the benchmark's roofline: an n-step forward
    Update is one links-only pass, n - 1 fused passes and one E-only pass, so
    it times n - 1 fused launches; a backward Update is n fused launches;
    runs past the reserved spans are counted as skipped; the timer must be
    started before it is stopped."""
    ext = (16, 16, 16)
    lattice(ext, SP)
    nrq.random_init(4)
    g = nrq.GaugeConfiguration()
    g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
    _say("fused-launch timer")
    nrq.leapfrog_timer_start(2)
    g.Update(1.0, 10)
    g.Update(-1.0, 4)
    g.Update(1.0, 5)  # past the two spans
    t = nrq.leapfrog_timer_stop()
    assert t["launches"] == 9 + 4 and t["skipped"] == 4
    assert t["site_steps"] == 13 * 16 ** 3
    assert t["ms"] > 0
    with pytest.raises(nrq.NRQError):
        nrq.leapfrog_timer_stop()
