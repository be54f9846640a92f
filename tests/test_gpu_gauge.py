"""GPU parity of the gauge half of the hot path against the CPU oracle.

Tolerances: FP64 links within 1e-12 relative after 10 leapfrog steps from an
identical initial condition; observables within 1e-12 relative; the initial
condition within 1e-13 absolute (hipFFT vs the oracle's FFT).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPACING_DT005 = (0.05, 1.0, 1.0, 1.0)


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_box_ic_matches_oracle(nrq, oracle, lattice):
    ext = (8, 8, 8)
    _say("box IC 8^3 seed 1")
    lattice(ext, SPACING_DT005)
    oracle.lattice_init(ext, SPACING_DT005)
    Uo, Eo = oracle.box_init(1, 0.5, 1.0, 1.0)
    nrq.random_init(1)
    g = nrq.GaugeConfiguration()
    g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
    U, E = g.download()
    assert np.max(np.abs(U - Uo)) < 1e-13
    assert np.max(np.abs(E - Eo)) < 1e-13


def test_box_ic_nonuniform_lattice(nrq, oracle, lattice):
    ext = (6, 4, 10)
    sp = (0.1, 1.0, 0.8, 1.3)
    _say("box IC 6x4x10 seed 7 anisotropic")
    lattice(ext, sp)
    oracle.lattice_init(ext, sp)
    Uo, Eo = oracle.box_init(7, 0.9, 1.5, 0.7)
    nrq.random_init(7)
    g = nrq.GaugeConfiguration()
    g.TransversePolarisedOccupiedInit_Box(0.9, 1.5, 0.7)
    U, E = g.download()
    assert np.max(np.abs(U - Uo)) < 1e-13
    assert np.max(np.abs(E - Eo)) < 1e-13


@pytest.mark.parametrize("ext", [(16, 16, 16), (5, 6, 7)])
def test_leapfrog_10_steps_fp64(nrq, oracle, lattice, ext):
    _say(f"leapfrog 10 steps {ext}")
    lattice(ext, SPACING_DT005)
    oracle.lattice_init(ext, SPACING_DT005)
    U, E = oracle.box_init(1, 0.5, 1.0, 1.0)
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    g.Update(1.0, 10)
    oracle.leapfrog(U, E, 1.0, 10)
    Ug, Eg = g.download()
    assert _rel(Ug, U) < 1e-12
    assert _rel(Eg, E) < 1e-12


def test_bench_workload_64cubed_matches_oracle(nrq, oracle, lattice):
    """The benchmarked configuration itself: 64^3, spacings (0.05, 1, 1, 1),
    box IC Qs = 0.5 from ranlxd seed 2.  The GPU's initial condition equals
    the oracle's (1e-13 absolute), 10 leapfrog steps from the same state
    agree to 1e-12 relative, and so do the observables the run logs."""
    ext = (64, 64, 64)
    _say("bench workload 64^3 seed 2: IC, 10 steps, observables")
    lattice(ext, SPACING_DT005)
    oracle.lattice_init(ext, SPACING_DT005)
    U, E = oracle.box_init(2, 0.5, 1.0, 1.0)
    nrq.random_init(2)
    g = nrq.GaugeConfiguration()
    g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
    Ug, Eg = g.download()
    assert np.max(np.abs(Ug - U)) < 1e-13
    assert np.max(np.abs(Eg - E)) < 1e-13
    g.upload(U, E)
    g.Update(1.0, 10)
    g.log_observables()
    oracle.leapfrog(U, E, 1.0, 10)
    Ug, Eg = g.download()
    assert _rel(Ug, U) < 1e-12
    assert _rel(Eg, E) < 1e-12
    row = g.read_log()[0]
    assert row[0] == pytest.approx(oracle.energy(U, E), rel=1e-12)
    assert row[2] == pytest.approx(oracle.gauss_max(U, E), rel=1e-12)
    assert row[3] == pytest.approx(oracle.mean_plaquette(U), rel=1e-12)


def test_leapfrog_single_calls_match_batched(nrq, oracle, lattice):
    ext = (8, 8, 8)
    _say("leapfrog single vs batched 8^3")
    lattice(ext, SPACING_DT005)
    oracle.lattice_init(ext, SPACING_DT005)
    U, E = oracle.box_init(2, 0.5, 1.0, 1.0)
    a = nrq.GaugeConfiguration()
    a.upload(U, E)
    b = a.copy()
    a.Update(1.0, 6)
    for _ in range(6):
        b.Update(1.0, 1)
    Ua, Ea = a.download()
    Ub, Eb = b.download()
    assert _rel(Ua, Ub) < 1e-14
    assert _rel(Ea, Eb) < 1e-14


def test_leapfrog_backward_step_matches_oracle(nrq, oracle, lattice):
    ext = (6, 6, 6)
    _say("leapfrog backward 6^3")
    lattice(ext, SPACING_DT005)
    oracle.lattice_init(ext, SPACING_DT005)
    U, E = oracle.box_init(3, 0.5, 1.0, 1.0)
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    g.Update(-1.0, 4)
    oracle.leapfrog(U, E, -1.0, 4)
    Ug, Eg = g.download()
    assert _rel(Ug, U) < 1e-12
    assert _rel(Eg, E) < 1e-12


def test_time_reversal_on_gpu(nrq, oracle, lattice):
    ext = (8, 8, 8)
    _say("time reversal 8^3")
    lattice(ext, SPACING_DT005)
    oracle.lattice_init(ext, SPACING_DT005)
    U, E = oracle.box_init(4, 0.5, 1.0, 1.0)
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    g.Update(1.0, 50)
    g.Update(-1.0, 50)
    Ug, Eg = g.download()
    assert np.max(np.abs(Ug - U)) < 1e-11
    assert np.max(np.abs(Eg - E)) < 1e-11


def test_observables_log_equals_synchronous_observables(nrq, lattice):
    """The device-side log holds exactly what nrq_gauge_observables returns
    at the same points (same kernels), across two growths of the log and a
    partial read."""
    ext = (6, 5, 4)
    _say("observables log 6x5x4, 150 rows")
    lattice(ext, SPACING_DT005)
    nrq.random_init(3)
    g = nrq.GaugeConfiguration()
    g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
    want, got = [], []
    for k in range(150):
        g.Update(1.0, 1 + k % 3)
        g.log_observables()
        o = g.observables()
        want.append([o["energy"], o["gauss_ref"], o["gauss_max"], o["plaquette"]])
        if k == 40:
            got.append(g.read_log(7))
            assert got[-1].shape == (7, 4)
    got.append(g.read_log())
    assert np.array_equal(np.concatenate(got), np.array(want))
    assert g.read_log().shape == (0, 4)


def test_observables_match_oracle(nrq, oracle, lattice):
    ext = (8, 6, 4)
    sp = (0.05, 1.0, 0.9, 1.1)
    _say("observables 8x6x4")
    lattice(ext, sp)
    oracle.lattice_init(ext, sp)
    U, E = oracle.box_init(5, 0.5, 1.0, 1.0)
    oracle.leapfrog(U, E, 1.0, 3)
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    obs = g.observables()
    assert obs["energy"] == pytest.approx(oracle.energy(U, E), rel=1e-12)
    assert obs["gauss_max"] == pytest.approx(oracle.gauss_max(U, E), rel=1e-12)
    assert obs["plaquette"] == pytest.approx(oracle.mean_plaquette(U), rel=1e-12)
    # the reference Gauss deviation is a sum of roundoff-level terms
    assert obs["gauss_ref"] < 1e-9 * obs["energy"]
    assert oracle.gauss_ref(U, E) < 1e-9 * obs["energy"]
    for (i, j, idx) in [(1, 2, 1), (2, 3, 17), (1, 3, 100), (3, 1, 55)]:
        assert np.allclose(g.GetPlaquette_G(i, j, idx), oracle.plaquette(U, i, j, idx), atol=1e-13)
    for (d, idx) in [(1, 1), (-1, 1), (2, 33), (-3, 190), (0, 5)]:
        ref = np.eye(3) if d == 0 else (U[:, :, d - 1, idx - 1] if d > 0 else
                                        U[:, :, -d - 1, oracle.neib(d, idx) - 1].conj().T)
        assert np.allclose(g.GetLink_G(d, idx), ref, atol=0)


def test_cold_start_is_fixed_point(nrq, lattice):
    _say("cold fixed point")
    lattice((4, 4, 4), SPACING_DT005)
    g = nrq.GaugeConfiguration()
    g.ColdInit()
    g.Update(1.0, 5)
    U, E = g.download()
    assert np.max(np.abs(E)) == 0.0
    eye = np.broadcast_to(np.eye(3)[:, :, None, None], U.shape)
    assert np.max(np.abs(U - eye)) == 0.0
    obs = g.observables()
    assert obs["energy"] == 0.0 and obs["plaquette"] == 1.0


def test_fp32_storage_tracks_fp64(nrq, oracle, lattice):
    ext = (8, 8, 8)
    _say("fp32 storage 8^3")
    lattice(ext, SPACING_DT005)
    oracle.lattice_init(ext, SPACING_DT005)
    U, E = oracle.box_init(6, 0.5, 1.0, 1.0)
    g = nrq.GaugeConfiguration(nrq.NRQ_PRECISION_FP32)
    g.upload(U, E)
    e0 = g.GetEnergy()
    g.Update(1.0, 10)
    g.Reunitarise()
    oracle.leapfrog(U, E, 1.0, 10)
    Ug, Eg = g.download()
    assert np.max(np.abs(Ug - U)) < 1e-4
    assert np.max(np.abs(Eg - E)) < 1e-4
    assert g.GetEnergy() == pytest.approx(oracle.energy(U, E), rel=1e-4)
    assert e0 == pytest.approx(oracle.energy(*oracle.box_init(6, 0.5, 1.0, 1.0)), rel=1e-5)


def test_mixed_precision_128cubed_errors(nrq, lattice):
    """The 128^3 mixed-precision workload (seed 3, Qs = 0.5): FP32 storage
    with FP64 reunitarisation every 10 steps against the all-FP64 run from the
    same seed, observables logged every 10 steps over 100 steps.  The bars are
    the errors measured over 500 steps (energy 9.1e-7, Gauss max 6.9e-7,
    plaquette 1.9e-5 relative) with a factor of about 2; the FP64 run conserves
    the Gauss law to roundoff."""
    ext = (128, 128, 128)
    _say("mixed precision 128^3 seed 3, 100 steps")
    lattice(ext, SPACING_DT005)
    logs = {}
    for prec in (nrq.NRQ_PRECISION_FP64, nrq.NRQ_PRECISION_FP32):
        nrq.random_init(3)
        g = nrq.GaugeConfiguration(prec)
        g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
        if prec == nrq.NRQ_PRECISION_FP32:
            g.SetReunitarisationInterval(10)
        g.log_observables()
        for _ in range(10):
            g.Update(1.0, 10)
            g.log_observables()
        logs[prec] = g.read_log()
        del g
    d, f = logs[nrq.NRQ_PRECISION_FP64], logs[nrq.NRQ_PRECISION_FP32]
    assert d.shape == f.shape == (11, 4)
    rel = np.abs(f - d) / np.abs(d)
    assert rel[:, 0].max() < 2e-6   # energy
    assert rel[:, 2].max() < 2e-6   # Gauss-law maximum
    assert rel[:, 3].max() < 4e-5   # mean plaquette
    assert np.abs(d[:, 2] - d[0, 2]).max() < 1e-12 * d[0, 2]


def test_reunitarisation_interval(nrq, oracle, lattice):
    """FP32 storage with automatic FP64 reunitarisation every 10 steps: the
    interval counts across Update calls and gives bit-identical results to
    explicit Reunitarise calls at the same points; the links stay unitary and
    track the FP64 oracle."""
    ext = (8, 6, 10)
    _say("reunitarisation interval")
    lattice(ext, SPACING_DT005)
    oracle.lattice_init(ext, SPACING_DT005)
    U, E = oracle.box_init(7, 0.5, 1.0, 1.0)
    auto = nrq.GaugeConfiguration(nrq.NRQ_PRECISION_FP32)
    auto.upload(U, E)
    auto.SetReunitarisationInterval(10)
    manual = auto.copy()
    manual.SetReunitarisationInterval(0)
    for n in (7, 8, 10):  # 25 steps: reunitarised after 10 and 20
        auto.Update(1.0, n)
    manual.Update(1.0, 10)
    manual.Reunitarise()
    manual.Update(1.0, 10)
    manual.Reunitarise()
    manual.Update(1.0, 5)
    Ua, Ea = auto.download()
    Um, Em = manual.download()
    assert np.array_equal(Ua, Um) and np.array_equal(Ea, Em)
    # unitarity of the stored FP32 links
    M = np.moveaxis(Ua.reshape(3, 3, 3, -1), (0, 1), (-2, -1))
    dev = np.abs(M @ np.conj(np.swapaxes(M, -1, -2)) - np.eye(3)).max()
    assert dev < 1e-5
    oracle.leapfrog(U, E, 1.0, 25)
    assert np.max(np.abs(Ua - U)) < 1e-4
    assert np.max(np.abs(Ea - E)) < 1e-4
    with pytest.raises(nrq.NRQError):
        auto.SetReunitarisationInterval(-1)


def test_reunitarisation_inside_the_lds_tile(nrq, oracle, lattice):
    """The LDS-tiled fused pass projects its links onto SU(3) at the
    reunitarisation points inside a call: with the tile shape pinned, calls of
    7, 8 and 10 steps with interval 10 agree with calls ending at the points
    plus explicit Reunitarise calls (to FP32 rounding: the half steps at the
    call boundaries are not the tile's code), the stored links are unitary,
    and both track the FP64 oracle."""
    ext = (32, 16, 16)
    _say("reunitarisation inside the LDS tile, 32x16x16")
    lattice(ext, SPACING_DT005)
    oracle.lattice_init(ext, SPACING_DT005)
    U, E = oracle.box_init(7, 0.5, 1.0, 1.0)
    tiles = [k for k, s in enumerate(nrq.leapfrog_shapes()) if s["tiling"].startswith("LDS") and s["applies_fp32"]]
    assert tiles
    nrq.leapfrog_set_shape(tiles[0])
    try:
        auto = nrq.GaugeConfiguration(nrq.NRQ_PRECISION_FP32)
        auto.upload(U, E)
        auto.SetReunitarisationInterval(10)
        manual = auto.copy()
        manual.SetReunitarisationInterval(0)
        for n in (7, 8, 10):  # 25 steps: reunitarised after 10 (inside a call) and 20 (inside a call)
            auto.Update(1.0, n)
        for n in (10, 10):
            manual.Update(1.0, n)
            manual.Reunitarise()
        manual.Update(1.0, 5)
        Ua, Ea = auto.download()
        Um, Em = manual.download()
    finally:
        nrq.leapfrog_set_shape(-1)
    assert np.max(np.abs(Ua - Um)) < 1e-5 and np.max(np.abs(Ea - Em)) < 1e-5 * max(1.0, np.max(np.abs(Em)))
    M = np.moveaxis(Ua.reshape(3, 3, 3, -1), (0, 1), (-2, -1))
    assert np.abs(M @ np.conj(np.swapaxes(M, -1, -2)) - np.eye(3)).max() < 1e-5
    oracle.leapfrog(U, E, 1.0, 25)
    assert np.max(np.abs(Ua - U)) < 1e-4
    assert np.max(np.abs(Ea - E)) < 1e-4 * max(1.0, np.max(np.abs(E)))


def test_ensemble_48cubed_batch_matches_single(nrq, lattice):
    """The ensemble workload's size: four FP32 48^3 members (seeds 100..103,
    Qs 0.3..0.6) with reunitarisation every 10 steps, 30 steps in one batched
    launch sequence per 10 steps (the hipGraph replay after the first), equal
    bit for bit to each member advanced on its own."""
    ext = (48, 48, 48)
    _say("ensemble 4 x 48^3 fp32 batched vs single")
    lattice(ext, SPACING_DT005)
    ens, ref = [], []
    for k in range(4):
        nrq.random_init(100 + k)
        g = nrq.GaugeConfiguration(nrq.NRQ_PRECISION_FP32)
        g.TransversePolarisedOccupiedInit_Box(0.3 + 0.1 * k, 1.0, 1.0)
        g.SetReunitarisationInterval(10)
        ens.append(g)
        ref.append(g.copy())
    for _ in range(3):
        nrq.leapfrog_batch(ens, 1.0, 10)
        for r in ref:
            r.Update(1.0, 10)
    for k, (a, b) in enumerate(zip(ens, ref)):
        Ua, Ea = a.download()
        Ub, Eb = b.download()
        assert np.array_equal(Ua, Ub), k
        assert np.array_equal(Ea, Eb), k


@pytest.mark.parametrize("interval", [0, 4])
@pytest.mark.parametrize("graph", ["1", "0"])
def test_leapfrog_batch_matches_single(nrq, lattice, monkeypatch, graph, interval):
    """An ensemble advanced by the batched launch (captured into a hipGraph and
    replayed on the repeated calls, or launched directly) equals each
    configuration advanced on its own, bit for bit, with and without a
    reunitarisation interval (the same for every member, so that both split
    the runs at the same steps) and for a backward step."""
    monkeypatch.setenv("NRQ_GRAPH", graph)
    ext = (8, 6, 6)
    _say(f"leapfrog batch graph={graph} interval={interval}")
    lattice(ext, SPACING_DT005)
    ens, ref = [], []
    for k in range(3):
        nrq.random_init(20 + k)
        g = nrq.GaugeConfiguration()
        g.TransversePolarisedOccupiedInit_Box(0.4 + 0.1 * k, 1.0, 1.0)
        g.SetReunitarisationInterval(interval)
        ens.append(g)
        ref.append(g.copy())
    for n, sw in ((3, 1.0), (6, 1.0), (6, 1.0), (2, -1.0)):
        nrq.leapfrog_batch(ens, sw, n)
        for r in ref:
            r.Update(sw, n)
    for k, (a, b) in enumerate(zip(ens, ref)):
        Ua, Ea = a.download()
        Ub, Eb = b.download()
        assert np.array_equal(Ua, Ub), k
        assert np.array_equal(Ea, Eb), k
    with pytest.raises(nrq.NRQError):
        nrq.leapfrog_batch([ens[0], ens[0]], 1.0, 1)


def test_leapfrog_batch_mixed_intervals(nrq, lattice):
    """Members with different reunitarisation intervals: the batch splits every
    member's run at each member's reunitarisation points, so a split member
    takes separate link and E half steps there instead of the fused pass; the
    result agrees with separate updates to rounding."""
    ext = (6, 6, 6)
    _say("leapfrog batch mixed intervals")
    lattice(ext, SPACING_DT005)
    ens, ref = [], []
    for k, iv in enumerate((0, 3, 5)):
        nrq.random_init(40 + k)
        g = nrq.GaugeConfiguration()
        g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
        g.SetReunitarisationInterval(iv)
        ens.append(g)
        ref.append(g.copy())
    nrq.leapfrog_batch(ens, 1.0, 11)
    for r in ref:
        r.Update(1.0, 11)
    for a, b in zip(ens, ref):
        Ua, Ea = a.download()
        Ub, Eb = b.download()
        assert np.max(np.abs(Ua - Ub)) < 1e-13
        assert np.max(np.abs(Ea - Eb)) < 1e-13
