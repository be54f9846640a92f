"""GPU parity and effect of the learned CN preconditioner (nrq_cnn.hip).

The checker is the numpy restatement oracle/cnn_oracle.py, which rounds to
bfloat16 where the GPU does; the GPU accumulates in FP32 (the oracle in
float64), so single hidden activations can round to the neighbouring bfloat16
value: the bars are relative L2 errors of 1e-2 (forward) and 3e-2 (gradient).
"""
import numpy as np
import pytest

from oracle import cnn_oracle as co

pytestmark = pytest.mark.gpu

SP = (0.05, 1.0, 1.0, 1.0)
C1 = [1.0, 0.0, 0.0, 0.0]


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("ext", [(16, 4, 3), (16, 5, 3), (32, 3, 5), (128, 2, 8), (128, 4, 16)])
def test_forward_matches_oracle(nrq, lattice, ext):
    """A wavefront convolves eight 16-site runs: 12 (the last wavefront has
    four runs), 15 (seven) and 30 runs (two per x row, six in the last); at
    x extent 128, even y and z a multiple of 8 a block sweeps 8 planes of a
    128 x 2 tile with its input in an LDS ring (halo wrapping in every
    direction; one and two blocks along z)."""
    _say(f"cnn forward {ext}")
    lattice(ext, SP)
    rng = np.random.default_rng(4)
    v = ext[0] * ext[1] * ext[2]
    inp = rng.normal(size=(v, co.CIN)).astype(np.float32)
    w = co.random_weights(rng).astype(np.float32)
    net = nrq.PreconditionerNet(1)
    net.weights = w
    assert np.array_equal(net.weights, w)
    out = net.forward(inp)
    ref, _, _ = co.forward(inp.reshape(ext[2], ext[1], ext[0], co.CIN), w)
    assert _rel(out, ref.reshape(v, co.CIN)) < 1e-2


@pytest.mark.parametrize("ext", [(16, 3, 2), (128, 2, 8)])
def test_loss_gradient_matches_oracle(nrq, lattice, ext):
    _say(f"cnn loss/gradient {ext}")
    lattice(ext, SP)
    rng = np.random.default_rng(5)
    v = ext[0] * ext[1] * ext[2]
    inp = rng.normal(size=(v, co.CIN)).astype(np.float32)
    tgt = rng.normal(size=(v, co.CIN)).astype(np.float32)
    w = co.random_weights(rng).astype(np.float32)
    net = nrq.PreconditionerNet(1)
    net.weights = w
    loss, grad = net.loss_grad(inp, tgt, 3.0)
    shp = (ext[2], ext[1], ext[0], co.CIN)
    rloss, rgrad = co.loss_grad(inp.reshape(shp), tgt.reshape(shp), w, 3.0)
    assert loss == pytest.approx(rloss, rel=1e-3)
    assert _rel(grad, rgrad) < 3e-2
    # per layer too, so that an error in one layer cannot hide behind the others
    g4, r4 = grad.reshape(4, -1), rgrad.reshape(4, -1)
    for l in range(4):
        assert _rel(g4[l], r4[l]) < 5e-2, l


@pytest.mark.parametrize("ext", [(16, 3, 2), (128, 2, 8)])
def test_each_layer_from_the_networks_own_values(nrq, lattice, ext):
    """Layer by layer from the GPU's own BF16 values, so that a rounding
    branch taken differently in float64 cannot hide or fake an error: the
    weight gradient of layer l is sum_s d_l(s) x act_l(s + tap) (float64 here)
    within 1e-3 relative; each layer input is the BF16 ReLU of the
    convolution of the previous one, each output gradient the BF16 ReLU-masked
    transposed convolution of the next, and the last one 2 (N - t) / norm
    (bars 1e-2: one BF16 rounding)."""
    _say(f"cnn layers from the GPU's own activations {ext}")
    lattice(ext, SP)
    rng = np.random.default_rng(6)
    v = ext[0] * ext[1] * ext[2]
    inp = rng.normal(size=(v, co.CIN)).astype(np.float32)
    tgt = rng.normal(size=(v, co.CIN)).astype(np.float32)
    w = co.random_weights(rng).astype(np.float32)
    net = nrq.PreconditionerNet(1)
    net.weights = w
    loss, grad, acts, deltas = net.loss_grad_trace(inp, tgt, 3.0)
    shp = (ext[2], ext[1], ext[0], co.CH)
    A = [a.reshape(shp).astype(np.float64) for a in acts]
    D = [d.reshape(shp).astype(np.float64) for d in deltas]
    ws = [co.bf16(x) for x in co.weights_layers(w)]
    first = np.zeros(shp)
    first[..., :co.CIN] = co.bf16(inp).reshape(shp[:3] + (co.CIN,))
    assert np.array_equal(A[0], first)
    g4 = grad.reshape(co.NL, co.TAPS, co.CH, co.CH)
    for l in range(co.NL):
        ref = np.stack([np.einsum("zyxo,zyxi->oi", D[l], co._shift(A[l], *off)) for off in co.OFFS])
        assert _rel(g4[l], ref) < 1e-3, l
    for l in range(co.NL - 1):
        assert _rel(A[l + 1], co.bf16(np.maximum(co.conv(A[l], ws[l]), 0.0))) < 1e-2, l
    for l in range(co.NL - 1, 0, -1):
        assert _rel(D[l - 1], co.bf16(np.where(A[l] > 0, co.conv_t(D[l], ws[l]), 0.0))) < 1e-2, l
    out = co.conv(A[co.NL - 1], ws[co.NL - 1])[..., :co.CIN]
    last = np.zeros(shp)
    last[..., :co.CIN] = co.bf16(2.0 * (out - tgt.reshape(shp[:3] + (co.CIN,))) / 3.0)
    assert _rel(D[co.NL - 1], last) < 1e-2
    assert loss == pytest.approx(float(np.sum((out - tgt.reshape(out.shape)) ** 2) / 3.0), rel=1e-3)


def _evolved(nrq, seed=7, steps=40):
    nrq.random_init(seed)
    g = nrq.GaugeConfiguration()
    g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
    g.Update(1.0, steps)
    return g


def test_fgmres_zero_network_is_the_neumann_polynomial(nrq, oracle, lattice):
    """With all weights zero the preconditioner is the second-order Neumann
    polynomial z = v - beta K v + (beta K)^2 v of (1 + beta K)^-1: FGMRES
    then needs fewer iterations than GMRES (the preconditioned operator is
    1 - (beta K)^3 + ..., |beta K| <= 0.075 here), and both reproduce the
    oracle's CN step."""
    ext = (16, 6, 5)
    _say("FGMRES zero network")
    lattice(ext, SP)
    nrq.nrqcd_init(2.0, C1, 1.0, 1e-12)
    try:
        oracle.lattice_init(ext, SP)
        U, E = oracle.box_init(8, 0.5, 1.0, 1.0)
        g = nrq.GaugeConfiguration()
        g.upload(U, E)
        net = nrq.PreconditionerNet(1)
        net.weights = np.zeros(nrq.NRQ_CNN_NWEIGHTS, np.float32)
        nrq.nrqcd_set_preconditioner(net)
        its = {}
        res = {}
        for name, kind in (("gmres", nrq.NRQ_SOLVER_GMRES_FP64), ("fgmres", nrq.NRQ_SOLVER_FGMRES_CNN)):
            nrq.nrqcd_set_solver(kind, 20, 10000)
            h = nrq.NRQCDField()
            h.InitSinglePoint(1, 17)
            st = h.Update(g, 2.0, C1)
            assert st.max_true_relres <= 1e-12
            its[name] = st.iterations
            res[name] = h.download()
        Q, A = oracle.point_prop(1), oracle.point_prop(17)
        oracle.cn_update(U, Q, A, 2.0, C1, 1.0, 1e-13)
        for name in res:
            assert np.max(np.abs(res[name][0] - Q)) < 1e-10 * np.max(np.abs(Q)), name
            assert np.max(np.abs(res[name][1] - A)) < 1e-10 * np.max(np.abs(A)), name
        _say(f"iterations: gmres {its['gmres']}, fgmres with the polynomial {its['fgmres']}")
        assert its["fgmres"] < its["gmres"]
        nrq.nrqcd_set_preconditioner(None)
    finally:
        nrq.nrqcd_finalize()


def test_training_and_gain(nrq, oracle, lattice):
    """Train on pairs from plain GMRES solves at Ma = 1, 2, 5 on an evolved
    configuration, then solve held-out masses 3 and 8 with FGMRES + the trained
    network, with FGMRES + the zero network (the Neumann polynomial alone)
    and with GMRES.  Shows what the network gains: the loss falls, all reach
    the same FP64 answer, and the trained network takes no more iterations
    per column solve than the polynomial alone, which takes fewer than GMRES
    (the counts are printed)."""
    ext = (16, 16, 8)
    _say("CNN training 16x16x8")
    lattice(ext, SP)
    nrq.nrqcd_init(2.0, C1, 1.0, 1e-10)
    try:
        g = _evolved(nrq)
        net = nrq.PreconditionerNet(3)
        for m in (1.0, 2.0, 5.0):
            net.harvest(g, m, 4, seed=int(10 * m))
        assert net.dataset_size() == 3 * 4 * 12
        first = net.train(1, 1e-3, seed=1)
        last = net.train(15, 1e-3, seed=2)
        _say(f"training loss: after 1 epoch {first:.4g}, after 16 epochs {last:.4g}")
        assert last < first
        zero = nrq.PreconditionerNet(3)
        zero.weights = np.zeros(nrq.NRQ_CNN_NWEIGHTS, np.float32)
        for m in (3.0, 8.0):
            out = {}
            for name, kind, pre in (("gmres", nrq.NRQ_SOLVER_GMRES_FP64, None),
                                    ("fgmres_polynomial", nrq.NRQ_SOLVER_FGMRES_CNN, zero),
                                    ("fgmres_cnn", nrq.NRQ_SOLVER_FGMRES_CNN, net),
                                    ("gmres_mixed", nrq.NRQ_SOLVER_GMRES_MIXED, None),
                                    ("fgmres_cnn_mixed", nrq.NRQ_SOLVER_FGMRES_CNN_MIXED, net)):
                nrq.nrqcd_set_preconditioner(pre)
                nrq.nrqcd_set_solver(kind, 20, 10000)
                h = nrq.NRQCDField()
                h.InitSinglePoint(100, 900)
                its = solves = 0
                for _ in range(3):
                    st = h.Update(g, m, C1)
                    assert st.max_true_relres <= 1e-10
                    its += st.iterations
                    solves += st.solves
                out[name] = (its / solves, h.download())
            _say(f"Ma={m}: iterations per column solve: " + ", ".join(f"{k} {v[0]:.2f}" for k, v in out.items()))
            assert out["fgmres_polynomial"][0] < out["gmres"][0]
            assert out["fgmres_cnn"][0] <= out["fgmres_polynomial"][0]
            for name in ("fgmres_polynomial", "fgmres_cnn", "gmres_mixed", "fgmres_cnn_mixed"):
                for k in range(2):
                    a, b = out[name][1][k], out["gmres"][1][k]
                    assert np.max(np.abs(a - b)) < 1e-8 * np.max(np.abs(b)), name
        nrq.nrqcd_set_preconditioner(None)
    finally:
        nrq.nrqcd_finalize()


@pytest.mark.parametrize("scale", [1e4, 30.0])
def test_fgmres_bad_network_falls_back_to_gmres(nrq, oracle, lattice, scale):
    """A network that does not precondition (weights far too large, so that
    N(v) swamps v in z = v + N(v)) must not change the
    answer: FGMRES hands the column solves it cannot finish to plain GMRES,
    which continues from the iterate reached and meets the oracle's CN step
    within 1e-10.  With the far too large weights the solves do fall back."""
    ext = (16, 6, 5)
    _say(f"FGMRES with a bad network (weights x {scale})")
    lattice(ext, SP)
    nrq.nrqcd_init(2.0, C1, 1.0, 1e-12)
    try:
        oracle.lattice_init(ext, SP)
        U, E = oracle.box_init(8, 0.5, 1.0, 1.0)
        g = nrq.GaugeConfiguration()
        g.upload(U, E)
        net = nrq.PreconditionerNet(5)
        net.weights = net.weights * np.float32(scale)
        nrq.nrqcd_set_preconditioner(net)
        nrq.nrqcd_set_solver(nrq.NRQ_SOLVER_FGMRES_CNN, 20, 10000)
        before = nrq.cn_fallbacks()
        h = nrq.NRQCDField()
        h.InitSinglePoint(1, 17)
        st = h.Update(g, 2.0, C1)
        fell = nrq.cn_fallbacks() - before
        _say(f"column solves handed to GMRES: {fell} of {st.solves}")
        assert st.max_true_relres <= 1e-12
        Q, A = oracle.point_prop(1), oracle.point_prop(17)
        oracle.cn_update(U, Q, A, 2.0, C1, 1.0, 1e-13)
        Qg, Ag = h.download()
        assert np.max(np.abs(Qg - Q)) < 1e-10 * np.max(np.abs(Q))
        assert np.max(np.abs(Ag - A)) < 1e-10 * np.max(np.abs(A))
        if scale > 1e3:
            assert fell > 0
        nrq.nrqcd_set_preconditioner(None)
    finally:
        nrq.nrqcd_finalize()


def test_learned_preconditioner_at_128cubed(nrq, lattice):
    """The learned-preconditioner configuration at its size: the 128^3 seed-3
    box initial condition evolved to t = 10/Qs (400 steps), pairs from plain
    GMRES solves at Ma = 1, 2, 5, the network trained on the GPU; at the
    held-out Ma = 3 and 8, FGMRES(20) with it (FP64, and FP32 inside FP64
    refinement) reaches FP64 true residual 1e-10 with the answer of
    GMRES(20), in fewer iterations per column solve (the iterations of the
    Neumann polynomial alone are printed beside them)."""
    ext = (128, 128, 128)
    _say("learned preconditioner at 128^3: evolve, harvest, train")
    lattice(ext, SP)
    nrq.nrqcd_init(2.0, C1, 1.0, 1e-10)
    try:
        g = _evolved(nrq, seed=3, steps=400)
        net = nrq.PreconditionerNet(3)
        for m in (1.0, 2.0, 5.0):
            net.harvest(g, m, 1, seed=int(10 * m))
        assert net.dataset_size() == 36
        first = net.train(1, 1e-3, seed=1)
        last = net.train(4, 1e-3, seed=2)
        _say(f"training loss: after 1 epoch {first:.4g}, after 5 epochs {last:.4g}")
        assert last < first
        zero = nrq.PreconditionerNet(3)
        zero.weights = np.zeros(nrq.NRQ_CNN_NWEIGHTS, np.float32)
        for m in (3.0, 8.0):
            out = {}
            for name, kind, pre in (("gmres", nrq.NRQ_SOLVER_GMRES_FP64, None),
                                    ("fgmres_polynomial", nrq.NRQ_SOLVER_FGMRES_CNN, zero),
                                    ("fgmres_cnn", nrq.NRQ_SOLVER_FGMRES_CNN, net),
                                    ("fgmres_cnn_mixed", nrq.NRQ_SOLVER_FGMRES_CNN_MIXED, net)):
                nrq.nrqcd_set_preconditioner(pre)
                nrq.nrqcd_set_solver(kind, 20, 10000)
                h = nrq.NRQCDField()
                h.InitSinglePoint(1, 1)
                st = h.Update(g, m, C1)
                assert st.max_true_relres <= 1e-10, name
                out[name] = (st.iterations / st.solves, h.download())
                h.Destructor()
            _say(f"128^3 Ma={m}: iterations per column solve: " + ", ".join(f"{k} {v[0]:.2f}" for k, v in out.items()))
            assert out["fgmres_cnn"][0] < out["gmres"][0]
            for name in ("fgmres_polynomial", "fgmres_cnn", "fgmres_cnn_mixed"):
                for k in range(2):
                    a, b = out[name][1][k], out["gmres"][1][k]
                    assert np.max(np.abs(a - b)) < 1e-9 * np.max(np.abs(b)), name
        nrq.nrqcd_set_preconditioner(None)
    finally:
        nrq.nrqcd_finalize()


def test_network_apply_timer(nrq, lattice):
    """nrq_cnn_time_forward times the device-resident apply; it leaves the
    weights alone and gives the same time scale for more applies."""
    lattice((32, 16, 16), SP)
    net = nrq.PreconditionerNet(5)
    w = net.weights.copy()
    a = net.time_forward(3)
    b = net.time_forward(12)
    _say(f"network apply at 32x16x16: {a:.4f} ms (3 applies), {b:.4f} ms (12 applies)")
    assert np.isfinite(a) and np.isfinite(b) and a > 0 and b > 0
    assert 0.2 < a / b < 5.0
    assert np.array_equal(net.weights, w)
    with pytest.raises(nrq.NRQError):
        net.time_forward(0)
