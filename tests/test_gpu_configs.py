"""The workload configurations at their full sizes on the GPU.

* 16^3 CPU configuration (seed 1, Ma = 2, point source at index 1, every
  column to relative residual 1e-12): CN then leapfrog per step, as the
  reference's driver orders them (main.f90:2329-2330), for 10 steps against
  the oracle (propagators within 1e-10, links within 1e-12), then continued to
  200 steps on the GPU with both norms conserved to the solver tolerance.
* 128^3 mixed-precision configuration (seed 3): the CN step with FP32 GMRES
  inside FP64 iterative refinement (and FP32 Chebyshev) against FP64 GMRES on
  the same FP64 configuration, all to FP64 true relative residual 1e-10.
* 32 x 48^3 ensemble (seeds 100..131, Qs uniform in [0.3, 0.7] from a
  separate generator, FP32 storage reunitarised every 10 steps): 30 steps in
  one batched, graph-replayed launch sequence equal bit for bit to single
  updates of members 0, 13 and 31; then the four masses 1, 2, 5, 10 at random
  sources of one member as one multi-right-hand-side CN solve (FP32
  Chebyshev corrections, FP64 outer residual 1e-10) against separate plain
  GMRES updates, norms conserved.
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


def test_cpu_configuration_16cubed_cn_then_leapfrog(nrq, oracle, lattice):
    ext = (16, 16, 16)
    lattice(ext, SP)
    oracle.lattice_init(ext, SP)
    nrq.nrqcd_init(2.0, C1, 1.0, 1e-12)
    try:
        nrq.nrqcd_set_solver(nrq.NRQ_SOLVER_GMRES_FP64, 30, 10000)
        U, E = oracle.box_init(1, 0.5, 1.0, 1.0)
        nrq.random_init(1)
        g = nrq.GaugeConfiguration()
        g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
        Ug, Eg = g.download()
        assert _rel(Ug, U) < 1e-13 and _rel(Eg, E) < 1e-13
        h = nrq.NRQCDField()
        h.InitSinglePoint(1, 1)
        Q, A = oracle.point_prop(1), oracle.point_prop(1)
        for step in range(1, 201):
            st = h.Update(g, 2.0, C1)
            assert st.max_true_relres <= 1e-12, step
            g.Update(1.0, 1)
            if step <= 10:
                rc, _ = oracle.cn_update(U, Q, A, 2.0, C1, 1.0, 1e-13)
                assert rc == 0
                oracle.leapfrog(U, E, 1.0, 1)
            if step == 10:
                _say("16^3 CN + leapfrog: 10 steps against the oracle")
                Qg, Ag = h.download()
                assert _rel(Qg, Q) < 1e-10 and _rel(Ag, A) < 1e-10
                Ug, Eg = g.download()
                assert _rel(Ug, U) < 1e-12 and _rel(Eg, E) < 1e-12
            if step % 50 == 0:
                nq, na = h.norms()
                _say(f"step {step}: norms - 1 = {nq - 1:.2e}, {na - 1:.2e}")
                # each solve changes the norm by at most ~ its residual 1e-12
                assert abs(nq - 1) < 1e-9 and abs(na - 1) < 1e-9, step
    finally:
        nrq.nrqcd_finalize()


def test_mixed_cn_128cubed_against_fp64_gmres(nrq, lattice):
    ext = (128, 128, 128)
    _say("CN at 128^3: FP32-GMRES / FP64-IR and FP32 Chebyshev against FP64 GMRES")
    lattice(ext, SP)
    nrq.nrqcd_init(2.0, C1, 1.0, 1e-10)
    try:
        nrq.random_init(3)
        g = nrq.GaugeConfiguration()
        g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
        g.Update(1.0, 10)
        out = {}
        for name in ("GMRES_FP64", "GMRES_MIXED", "CHEBYSHEV_MIXED"):
            nrq.nrqcd_set_solver(getattr(nrq, "NRQ_SOLVER_" + name), 30, 10000)
            h = nrq.NRQCDField()
            h.InitSinglePoint(1, 1)
            for _ in range(2):
                st = h.Update(g, 2.0, C1)
                assert st.max_true_relres <= 1e-10, name
            nq, na = h.norms()
            assert abs(nq - 1) < 1e-9 and abs(na - 1) < 1e-9, name
            out[name] = h.download()
            h.Destructor()
        Q0, A0 = out["GMRES_FP64"]
        for name in ("GMRES_MIXED", "CHEBYSHEV_MIXED"):
            Q, A = out[name]
            _say(f"{name}: max difference to FP64 GMRES {max(_rel(Q, Q0), _rel(A, A0)):.2e}")
            # both solutions are within ~cond * 1e-10 of the exact one
            assert _rel(Q, Q0) < 1e-9 and _rel(A, A0) < 1e-9, name
    finally:
        nrq.nrqcd_finalize()


def test_ensemble_32x48cubed_full_size(nrq, lattice, monkeypatch):
    monkeypatch.setenv("NRQ_GRAPH", "1")
    ext = (48, 48, 48)
    lattice(ext, SP)
    vol = 48 ** 3
    rng = np.random.default_rng(1000)  # Qs and sources, apart from the ranlxd stream of each seed
    qs = rng.uniform(0.3, 0.7, 32)
    _say("ensemble 32 x 48^3 fp32: initial conditions")
    ens = []
    for k in range(32):
        nrq.random_init(100 + k)
        g = nrq.GaugeConfiguration(nrq.NRQ_PRECISION_FP32)
        g.TransversePolarisedOccupiedInit_Box(float(qs[k]), 1.0, 1.0)
        g.SetReunitarisationInterval(10)
        ens.append(g)
    check = (0, 13, 31)
    ref = {k: ens[k].copy() for k in check}
    _say("ensemble 32 x 48^3 fp32: 30 batched steps against single updates")
    for _ in range(3):
        nrq.leapfrog_batch(ens, 1.0, 10)
        for r in ref.values():
            r.Update(1.0, 10)
    for k in check:
        Ua, Ea = ens[k].download()
        Ub, Eb = ref[k].download()
        assert np.array_equal(Ua, Ub) and np.array_equal(Ea, Eb), k
    e = [g.GetEnergy() for g in ens]
    assert all(np.isfinite(e)) and len(set(e)) == 32
    del ref

    _say("ensemble member 13: four masses as one multi-right-hand-side CN solve")
    nrq.nrqcd_init(2.0, C1, 1.0, 1e-10)
    try:
        masses = [1.0, 2.0, 5.0, 10.0]
        sites = rng.integers(1, vol + 1, size=(4, 2))
        batch, single = [], []
        for iq, ia in sites:
            h = nrq.NRQCDField()
            h.InitSinglePoint(int(iq), int(ia))
            batch.append(h)
            single.append(h.copy())
        nrq.nrqcd_set_solver(nrq.NRQ_SOLVER_CHEBYSHEV_MIXED, 30, 10000)
        for _ in range(2):
            st = nrq.update_batch(batch, ens[13], masses, C1)
            assert st.solves == 48 and st.max_true_relres <= 1e-10
        nrq.nrqcd_set_solver(nrq.NRQ_SOLVER_GMRES_FP64, 30, 10000)
        for _ in range(2):
            for h, m in zip(single, masses):
                assert h.Update(ens[13], m, C1).max_true_relres <= 1e-10
        for m, hb, hs in zip(masses, batch, single):
            Qb, Ab = hb.download()
            Qs, As = hs.download()
            assert _rel(Qb, Qs) < 1e-9 and _rel(Ab, As) < 1e-9, m
            nq, na = hb.norms()
            assert abs(nq - 1) < 1e-9 and abs(na - 1) < 1e-9, m
    finally:
        nrq.nrqcd_finalize()
