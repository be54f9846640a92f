"""GPU parity of the NRQCD Crank-Nicolson half of the hot path.

Both sides solve to an FP64 true relative residual of 1e-12; the linear
solution is unique, so the propagators must agree to ~cond * 1e-12 (the
system is within 0.3% of the identity here): tolerance 1e-10 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SP = (0.05, 1.0, 1.0, 1.0)
C1 = [1.0, 0.0, 0.0, 0.0]


def _say(msg):
    print(msg, flush=True)


@pytest.fixture()
def cn(nrq, lattice):
    """Factory: lattice::InitModule, then nrqcd::InitModule (the reference's
    order, nrqcd.f90:167-192 checks the lattice module first)."""
    state = {"on": False}

    def init(ext, spacing, tol=1e-12):
        lattice(ext, spacing)
        nrq.nrqcd_init(2.0, C1, 1.0, tol)
        state["on"] = True

    yield init
    if state["on"]:
        nrq.nrqcd_finalize()


def test_cn_step_matches_oracle(nrq, oracle, cn):
    ext = (8, 8, 8)
    _say("CN 3 steps 8^3")
    cn(ext, SP)
    oracle.lattice_init(ext, SP)
    U, E = oracle.box_init(1, 0.5, 1.0, 1.0)
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    h = nrq.NRQCDField()
    h.InitSinglePoint(1, 1)
    Q = oracle.point_prop(1)
    A = oracle.point_prop(1)
    for step in range(3):
        st = h.Update(g, 2.0, C1)
        rc, ost = oracle.cn_update(U, Q, A, 2.0, C1, 1.0, 1e-12)
        assert rc == 0
        assert st.solves == 12
        assert st.max_true_relres <= 1e-12
    Qg, Ag = h.download()
    scale = np.max(np.abs(Q))
    assert np.max(np.abs(Qg - Q)) < 1e-10 * scale
    assert np.max(np.abs(Ag - A)) < 1e-10 * scale
    nq, na = h.norms()
    assert nq == pytest.approx(oracle.prop_norm(Q), rel=1e-11)
    assert na == pytest.approx(oracle.prop_norm(A), rel=1e-11)
    for gconjg in (False, True):
        corr = h.correlators(gconjg)
        ocorr = oracle.correlators(Q, A, gconjg)
        for k in ocorr:
            assert abs(corr[k] - ocorr[k]) < 1e-10 * max(1.0, abs(ocorr[k])), (k, gconjg)


def test_cn_general_propagator_and_anisotropy(nrq, oracle, cn):
    ext = (4, 6, 5)
    sp = (0.1, 1.0, 0.7, 1.2)
    _say("CN random propagator 4x6x5")
    cn(ext, sp)
    oracle.lattice_init(ext, sp)
    U, E = oracle.box_init(9, 0.7, 1.0, 1.0)
    rng = np.random.default_rng(3)
    v = oracle.vol
    Q = np.asfortranarray(rng.normal(size=(6, 6, v)) + 1j * rng.normal(size=(6, 6, v)))
    A = np.asfortranarray(rng.normal(size=(6, 6, v)) + 1j * rng.normal(size=(6, 6, v)))
    c = [0.8 + 0.1j, 0, 0, 0]
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    h = nrq.NRQCDField()
    h.upload(Q, A)
    Qd, Ad = h.download()
    assert np.array_equal(Qd, Q) and np.array_equal(Ad, A)
    # all eight correlators of general (non-colour-diagonal) propagators:
    # the GPU's trace form against the oracle's explicit vertex sums
    for gconjg in (False, True):
        corr = h.correlators(gconjg)
        ocorr = oracle.correlators(Q, A, gconjg)
        for k in ocorr:
            assert abs(corr[k] - ocorr[k]) < 1e-12 * abs(ocorr[k]), (k, gconjg)
    h.Update(g, 1.3, c, 0.5)
    rc, _ = oracle.cn_update(U, Q, A, 1.3, c, 0.5, 1e-12)
    assert rc == 0
    Qg, Ag = h.download()
    assert np.max(np.abs(Qg - Q)) < 1e-10 * np.max(np.abs(Q))
    assert np.max(np.abs(Ag - A)) < 1e-10 * np.max(np.abs(A))


def _observables(h):
    return (h.correlators(), h.correlators(True), h.norms())


def _oracle_observables(oracle, Q, A):
    return (oracle.correlators(Q, A), oracle.correlators(Q, A, True), (oracle.prop_norm(Q), oracle.prop_norm(A)))


def _same(got, want, rel=1e-11):
    for g, w in zip(got[:2], want[:2]):
        for k in w:
            assert abs(g[k] - w[k]) <= rel * max(1.0, abs(w[k])), k
    assert got[2][0] == pytest.approx(want[2][0], rel=rel, abs=1e-300)
    assert got[2][1] == pytest.approx(want[2][1], rel=rel, abs=1e-300)


def test_observables_follow_every_change(nrq, oracle, cn):
    """The library evaluates the eight correlators and both norms in one pass
    and reuses them until the field changes: after creation, point sources,
    uploads, copies and CN steps the values must be those of the new field."""
    ext = (6, 5, 4)
    _say("observables after every change 6x5x4")
    cn(ext, SP)
    oracle.lattice_init(ext, SP)
    U, E = oracle.box_init(4, 0.5, 1.0, 1.0)
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    h = nrq.NRQCDField()
    z = _observables(h)
    assert z[2] == (0.0, 0.0) and all(v == 0 for v in z[0].values())
    h.InitSinglePoint(1, 1)
    Q, A = oracle.point_prop(1), oracle.point_prop(1)
    _same(_observables(h), _oracle_observables(oracle, Q, A))
    h.Update(g, 2.0, C1)
    rc, _ = oracle.cn_update(U, Q, A, 2.0, C1, 1.0, 1e-12)
    assert rc == 0
    _same(_observables(h), _oracle_observables(oracle, Q, A), rel=1e-10)
    c = h.copy()
    _same(_observables(c), _observables(h), rel=0)
    c.Update(g, 2.0, C1)
    rc, _ = oracle.cn_update(U, Q, A, 2.0, C1, 1.0, 1e-12)
    _same(_observables(c), _oracle_observables(oracle, Q, A), rel=1e-10)
    assert abs(c.correlators()["O1_1S0"] - h.correlators()["O1_1S0"]) > 1e-6
    rng = np.random.default_rng(8)
    v = oracle.vol
    Qr = np.asfortranarray(rng.normal(size=(6, 6, v)) + 1j * rng.normal(size=(6, 6, v)))
    h.upload(Qr, None)
    _same(_observables(h), _oracle_observables(oracle, Qr, h.download()[1]), rel=1e-12)
    h.InitSinglePointSingleDoF(1, 2, 3, 2, 1, 3)
    Qd, Ad = h.download()
    assert np.count_nonzero(Qd) == 1 and np.count_nonzero(Ad) == 1
    _same(_observables(h), _oracle_observables(oracle, Qd, Ad), rel=1e-12)


def test_cn_norm_conservation_and_reversibility(nrq, cn):
    ext = (8, 8, 8)
    _say("CN reversibility 8^3")
    cn(ext, SP)
    nrq.random_init(11)
    g = nrq.GaugeConfiguration()
    g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
    h = nrq.NRQCDField()
    h.InitSinglePoint(1, 1)
    Q0, A0 = h.download()
    for _ in range(5):
        h.Update(g, 2.0, C1)
    nq, na = h.norms()
    assert abs(nq - 1.0) < 1e-10 and abs(na - 1.0) < 1e-10
    for _ in range(5):
        h.Update(g, 2.0, C1, -1.0)
    Q1, A1 = h.download()
    assert np.max(np.abs(Q1 - Q0)) < 1e-10
    assert np.max(np.abs(A1 - A0)) < 1e-10


def test_bench_workload_64cubed_cn_properties(nrq, cn):
    """At the benchmarked size (64^3, box IC seed 2 evolved 20 steps, point
    sources, Ma = 2, FP64 true residual 1e-10) with the fastest solver,
    mixed-precision Chebyshev: both norms stay 1 to the solver tolerance over
    3 steps, the steps match the plain FP64 GMRES within the tolerance, and 3
    backward steps (StepWidth -1, the exact inverse of a CN step) return the
    point sources."""
    ext = (64, 64, 64)
    _say("CN 64^3 mixed Chebyshev: norms, against GMRES, reversibility")
    cn(ext, (0.05, 1.0, 1.0, 1.0), tol=1e-10)
    nrq.random_init(2)
    g = nrq.GaugeConfiguration()
    g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
    g.Update(1.0, 20)
    fields = {}
    for kind in (nrq.NRQ_SOLVER_GMRES_FP64, nrq.NRQ_SOLVER_CHEBYSHEV_MIXED):
        nrq.nrqcd_set_solver(kind, 30, 10000)
        h = nrq.NRQCDField()
        h.InitSinglePoint(1, 1)
        for _ in range(3):
            st = h.Update(g, 2.0, C1)
            assert st.max_true_relres <= 1e-10
        nq, na = h.norms()
        assert abs(nq - 1.0) < 1e-9 and abs(na - 1.0) < 1e-9
        fields[kind] = (h, h.download())
    (_, (Qg, Ag)), (h, (Qc, Ac)) = fields[nrq.NRQ_SOLVER_GMRES_FP64], fields[nrq.NRQ_SOLVER_CHEBYSHEV_MIXED]
    assert np.max(np.abs(Qc - Qg)) < 1e-9 and np.max(np.abs(Ac - Ag)) < 1e-9
    for _ in range(3):
        h.Update(g, 2.0, C1, -1.0)
    Q1, A1 = h.download()
    h0 = nrq.NRQCDField()
    h0.InitSinglePoint(1, 1)
    Q0, A0 = h0.download()
    assert np.max(np.abs(Q1 - Q0)) < 1e-9 and np.max(np.abs(A1 - A0)) < 1e-9


def test_init_single_dof_matches_reference_indexing(nrq, lattice):
    lattice((4, 4, 4), SP)
    h = nrq.NRQCDField()
    h.InitSinglePointSingleDoF(2, 3, 5, 1, 2, 7)
    Q, A = h.download()
    # nrqcd.f90:437-444: index = colour + 3 (spin - 1)
    iq = 3 + 3 * (2 - 1)
    ia = 2 + 3 * (1 - 1)
    assert Q[iq - 1, iq - 1, 5 - 1] == 1.0 and np.count_nonzero(Q) == 1
    assert A[ia - 1, ia - 1, 7 - 1] == 1.0 and np.count_nonzero(A) == 1


SOLVERS = ["GMRES_MIXED", "CHEBYSHEV_FP64", "CHEBYSHEV_MIXED"]


@pytest.mark.parametrize("kind", SOLVERS)
@pytest.mark.parametrize("tol", [1e-12, 1e-10])
def test_fast_solvers_match_oracle(nrq, oracle, cn, tol, kind):
    """Mixed-precision GMRES and the Chebyshev solvers: the FP64 true
    residual decides convergence, so each result agrees with the FP64 oracle
    to the same bar as the all-FP64 GMRES."""
    ext = (8, 6, 4)
    sp = (0.05, 1.0, 0.9, 1.1)
    _say(f"CN {kind} tol {tol}")
    cn(ext, sp, tol)
    nrq.nrqcd_set_solver(getattr(nrq, "NRQ_SOLVER_" + kind), 30, 10000)
    oracle.lattice_init(ext, sp)
    U, E = oracle.box_init(3, 0.5, 1.0, 1.0)
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    h = nrq.NRQCDField()
    h.InitSinglePoint(1, 7)
    Q, A = oracle.point_prop(1), oracle.point_prop(7)
    for _ in range(3):
        st = h.Update(g, 2.0, C1)
        assert st.max_true_relres <= tol
        assert st.solves == 12
        oracle.cn_update(U, Q, A, 2.0, C1, 1.0, 1e-13)
    Qg, Ag = h.download()
    bar = 200 * tol
    assert np.max(np.abs(Qg - Q)) < bar * np.max(np.abs(Q))
    assert np.max(np.abs(Ag - A)) < bar * np.max(np.abs(A))


@pytest.mark.parametrize("kind", SOLVERS)
def test_fast_solvers_on_fp32_gauge_storage(nrq, oracle, cn, kind):
    """FP32 link storage: the FP64 outer residual uses the stored links, so
    the solve is exact for that operator; compare with the oracle on the
    FP32-rounded links."""
    ext = (6, 6, 6)
    _say(f"CN {kind} on FP32 links")
    cn(ext, SP, 1e-11)
    nrq.nrqcd_set_solver(getattr(nrq, "NRQ_SOLVER_" + kind), 30, 10000)
    oracle.lattice_init(ext, SP)
    U, E = oracle.box_init(4, 0.5, 1.0, 1.0)
    U32 = np.asfortranarray(U.astype(np.complex64).astype(np.complex128))
    g = nrq.GaugeConfiguration(nrq.NRQ_PRECISION_FP32)
    g.upload(U, E)
    h = nrq.NRQCDField()
    h.InitSinglePoint(3, 3)
    st = h.Update(g, 2.0, C1)
    assert st.max_true_relres <= 1e-11
    Q, A = oracle.point_prop(3), oracle.point_prop(3)
    oracle.cn_update(U32, Q, A, 2.0, C1, 1.0, 1e-13)
    Qg, Ag = h.download()
    assert np.max(np.abs(Qg - Q)) < 5e-9 * np.max(np.abs(Q))


def test_chebyshev_complex_coefficient_and_anisotropy(nrq, oracle, cn):
    """Chebyshev segment for complex c1 and anisotropic spacings, and the
    reverse step (StepWidth < 0 flips the segment)."""
    ext = (4, 6, 5)
    sp = (0.1, 1.0, 0.7, 1.2)
    _say("CN Chebyshev complex c1")
    cn(ext, sp, 1e-12)
    nrq.nrqcd_set_solver(nrq.NRQ_SOLVER_CHEBYSHEV_FP64, 30, 10000)
    oracle.lattice_init(ext, sp)
    U, E = oracle.box_init(9, 0.7, 1.0, 1.0)
    rng = np.random.default_rng(5)
    v = oracle.vol
    Q = np.asfortranarray(rng.normal(size=(6, 6, v)) + 1j * rng.normal(size=(6, 6, v)))
    A = np.asfortranarray(rng.normal(size=(6, 6, v)) + 1j * rng.normal(size=(6, 6, v)))
    c = [0.8 + 0.3j, 0, 0, 0]
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    h = nrq.NRQCDField()
    h.upload(Q, A)
    for sw in (0.5, -1.0):
        st = h.Update(g, 1.3, c, sw)
        assert st.max_true_relres <= 1e-12
        rc, _ = oracle.cn_update(U, Q, A, 1.3, c, sw, 1e-13)
        assert rc == 0
    Qg, Ag = h.download()
    assert np.max(np.abs(Qg - Q)) < 2e-10 * np.max(np.abs(Q))
    assert np.max(np.abs(Ag - A)) < 2e-10 * np.max(np.abs(A))


@pytest.mark.parametrize("kind", ["GMRES_FP64"] + SOLVERS)
def test_partially_zero_propagators(nrq, oracle, cn, kind):
    """Only the live (non-zero) fields enter the batched solve: a quark
    propagator with two zero columns and a zero antiquark propagator evolve
    like the oracle's, and the zero parts stay exactly zero.  A copy keeps the
    same set of live fields."""
    ext = (6, 4, 5)
    _say(f"CN partial propagator {kind}")
    cn(ext, SP, 1e-12)
    nrq.nrqcd_set_solver(getattr(nrq, "NRQ_SOLVER_" + kind), 30, 10000)
    oracle.lattice_init(ext, SP)
    U, E = oracle.box_init(5, 0.5, 1.0, 1.0)
    rng = np.random.default_rng(8)
    v = oracle.vol
    Q = np.asfortranarray(rng.normal(size=(6, 6, v)) + 1j * rng.normal(size=(6, 6, v)))
    Q[:, 1, :] = 0
    Q[:, 4, :] = 0
    A = np.zeros((6, 6, v), dtype=np.complex128, order="F")
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    h = nrq.NRQCDField()
    h.upload(Q, A)
    h2 = h.copy()
    for _ in range(2):
        st = h.Update(g, 2.0, C1)
        assert st.solves == 8  # 4 non-zero columns x 2 spins
        assert st.max_true_relres <= 1e-12
        h2.Update(g, 2.0, C1)
        oracle.cn_update(U, Q, A, 2.0, C1, 1.0, 1e-13)
    Qg, Ag = h.download()
    Q2, A2 = h2.download()
    assert np.array_equal(Qg, Q2) and np.array_equal(Ag, A2)
    assert not np.any(Qg[:, 1, :]) and not np.any(Qg[:, 4, :]) and not np.any(Ag)
    assert np.max(np.abs(Qg - Q)) < 2e-10 * np.max(np.abs(Q))


@pytest.mark.parametrize("kind", ["CHEBYSHEV_FP64", "CHEBYSHEV_MIXED"])
def test_chebyshev_long_step(nrq, oracle, cn, kind):
    """A step 40x longer than the default (|beta| kmax = 3): the a priori
    plan no longer suffices and the measured refinement rounds finish the
    solve to the same FP64 bar."""
    ext = (6, 6, 6)
    _say(f"CN {kind} long step")
    cn(ext, SP, 1e-12)
    nrq.nrqcd_set_solver(getattr(nrq, "NRQ_SOLVER_" + kind), 30, 10000)
    oracle.lattice_init(ext, SP)
    U, E = oracle.box_init(6, 0.5, 1.0, 1.0)
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    h = nrq.NRQCDField()
    h.InitSinglePoint(5, 9)
    Q, A = oracle.point_prop(5), oracle.point_prop(9)
    for _ in range(2):
        st = h.Update(g, 2.0, C1, 40.0)
        assert st.max_true_relres <= 1e-12
        rc, _ = oracle.cn_update(U, Q, A, 2.0, C1, 40.0, 1e-13)
        assert rc == 0
    Qg, Ag = h.download()
    assert np.max(np.abs(Qg - Q)) < 1e-9 * np.max(np.abs(Q))
    assert np.max(np.abs(Ag - A)) < 1e-9 * np.max(np.abs(A))


@pytest.mark.parametrize("kind", ["GMRES_FP64", "CHEBYSHEV_MIXED"])
def test_multi_mass_batch_matches_single_updates(nrq, oracle, cn, kind):
    """Four heavy fields with masses 1, 2, 5 and 10 at different sources on one
    configuration, advanced as one 48-right-hand-side solve, agree with
    separate updates and with the oracle."""
    ext = (6, 6, 4)
    _say(f"CN multi-mass batch {kind}")
    cn(ext, SP, 1e-12)
    nrq.nrqcd_set_solver(getattr(nrq, "NRQ_SOLVER_" + kind), 30, 10000)
    oracle.lattice_init(ext, SP)
    U, E = oracle.box_init(12, 0.5, 1.0, 1.0)
    g = nrq.GaugeConfiguration()
    g.upload(U, E)
    masses = [1.0, 2.0, 5.0, 10.0]
    sites = [(1, 5), (9, 30), (77, 2), (144, 100)]
    batch, single = [], []
    for (iq, ia) in sites:
        h = nrq.NRQCDField()
        h.InitSinglePoint(iq, ia)
        batch.append(h)
        single.append(h.copy())
    for _ in range(2):
        st = nrq.update_batch(batch, g, masses, C1)
        assert st.solves == 48 and st.max_true_relres <= 1e-12
        for h, m in zip(single, masses):
            h.Update(g, m, C1)
    for (iq, ia), m, hb, hs in zip(sites, masses, batch, single):
        Qb, Ab = hb.download()
        Qs, As = hs.download()
        assert np.max(np.abs(Qb - Qs)) < 1e-10 * np.max(np.abs(Qs))
        Q, A = oracle.point_prop(iq), oracle.point_prop(ia)
        for _ in range(2):
            oracle.cn_update(U, Q, A, m, C1, 1.0, 1e-13)
        assert np.max(np.abs(Qb - Q)) < 1e-10 * np.max(np.abs(Q))
        assert np.max(np.abs(Ab - A)) < 1e-10 * np.max(np.abs(A))
