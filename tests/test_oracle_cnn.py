"""Pinning the CPU restatement of the preconditioner network (no GPU)."""
import numpy as np

from oracle import cnn_oracle as co


def test_conv_matches_site_loop():
    """The vectorised periodic convolution equals the definition evaluated
    site by site on a tiny anisotropic lattice."""
    rng = np.random.default_rng(1)
    n2, n1, n0 = 2, 3, 4
    x = rng.normal(size=(n2, n1, n0, 5))
    w = rng.normal(size=(27, 4, 5))
    out = co.conv(x, w)
    ref = np.zeros_like(out)
    for z in range(n2):
        for y in range(n1):
            for xx in range(n0):
                for t, (dx, dy, dz) in enumerate(co.OFFS):
                    ref[z, y, xx] += w[t] @ x[(z + dz) % n2, (y + dy) % n1, (xx + dx) % n0]
    assert np.allclose(out, ref, rtol=1e-13, atol=1e-13)


def test_conv_t_is_adjoint():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(3, 4, 5, 6))
    d = rng.normal(size=(3, 4, 5, 7))
    w = rng.normal(size=(27, 7, 6))
    assert np.isclose(np.sum(co.conv(x, w) * d), np.sum(x * co.conv_t(d, w)), rtol=1e-12)


def test_gradient_matches_finite_differences():
    """Analytic backward pass (no bfloat16 rounding) against central
    differences of the loss for a sample of weights in every layer."""
    rng = np.random.default_rng(3)
    shape = (2, 2, 4)
    inp = rng.normal(size=shape + (co.CIN,))
    tgt = rng.normal(size=shape + (co.CIN,))
    w = co.random_weights(rng)
    loss, g = co.loss_grad(inp, tgt, w, 7.0, bf=False)
    layer = co.TAPS * co.CH * co.CH
    picks = []
    for l in range(co.NL):
        cin = co.CIN if l == 0 else co.CH
        cout = co.CIN if l == co.NL - 1 else co.CH
        for _ in range(6):
            t, o, i = rng.integers(27), rng.integers(cout), rng.integers(cin)
            picks.append(l * layer + (t * co.CH + o) * co.CH + i)
    h = 1e-6
    for k in picks:
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (co.loss_grad(inp, tgt, wp, 7.0, bf=False)[0] - co.loss_grad(inp, tgt, wm, 7.0, bf=False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
    # padded weights get no gradient
    g4 = g.reshape(co.NL, co.TAPS, co.CH, co.CH)
    assert not np.any(g4[0, :, :, co.CIN:]) and not np.any(g4[3, :, co.CIN:, :])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 0.0, 1e-30], dtype=np.float32)
    r = co.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7  # ties to even
    assert r[3] == -2.5 and r[4] == 0.0
    assert abs(r[5] - 1e-30) < 1e-32
