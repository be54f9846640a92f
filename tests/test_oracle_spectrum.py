"""Pinning the CPU restatement of the spectral-function post-processing
(oracle/spectrum_oracle.py, main.f90:1865-2100); no GPU."""
import numpy as np
import pytest

from oracle import spectrum_oracle as so


def _geometric(q, lo, hi):
    """sum_{m=lo}^{hi} q^m"""
    q = np.asarray(q, np.complex128)
    cnt = hi - lo + 1
    out = np.empty_like(q)
    near = np.abs(q - 1) < 1e-13
    out[near] = cnt
    qq = q[~near]
    out[~near] = qq ** lo * (1 - qq ** cnt) / (1 - qq)
    return out


@pytest.mark.parametrize("n, dt, E, amp", [(37, 0.1, 1.2345, 0.7 + 0.2j), (64, 0.05, -3.1, 1.0), (2, 0.3, 0.9, 1j)])
def test_single_exponential_closed_form(n, dt, E, amp):
    """C(t) = a exp(-i E t).  The symmetrised, windowed signal at time m dt
    (m = -(n-1) .. n-2) is w(m) x(m) with x(m) = a z^m for m >= 0 and
    -conj(a) z^m for m < 0 (z = exp(-i E dt)), and the Hann window centred
    on t = 0 is w(m) = cos^2(pi m / N) = 1/2 + (e^{2 pi i m/N} + e^{-2 pi i m/N})/4.
    The backward transform is then a sum of six geometric series."""
    t = np.arange(n) * dt
    c = amp * np.exp(-1j * E * t)
    spec, _ = so.spectrum(c, dt)
    N = 2 * n - 2
    ell = np.arange(N)
    F = np.zeros(N, np.complex128)
    for coef, phi in ((0.5, 0.0), (0.25, 2 * np.pi / N), (0.25, -2 * np.pi / N)):
        q = np.exp(1j * (2 * np.pi * ell / N - E * dt + phi))
        F += coef * (amp * _geometric(q, 0, n - 2) - np.conj(amp) * _geometric(q, -(n - 1), -1))
    want = 2 * np.imag(dt * F)
    assert np.allclose(spec, want, rtol=0, atol=1e-12 * max(1.0, np.abs(want).max()))


def test_transform_real_part_is_constant():
    """Negative times carry -conj of the positive ones and the window vanishes
    at the one unpaired time, so the transform is i (real) + dt Re C(0)."""
    rng = np.random.default_rng(5)
    c = rng.normal(size=50) + 1j * rng.normal(size=50)
    dt = 0.07
    _, (_, _, f) = so.spectrum(c, dt)
    F = so.t2w(f, dt)
    assert np.allclose(F.real, dt * c[0].real, atol=1e-12)


def test_stages():
    c = np.arange(1, 6) * (1 + 2j)  # n = 5 -> N = 8
    _, (s, h, f) = so.spectrum(c, 1.0)
    want_s = [-np.conj(c[4]), -np.conj(c[3]), -np.conj(c[2]), -np.conj(c[1]), c[0], c[1], c[2], c[3]]
    assert np.array_equal(s, want_s)
    assert h[0] == 0 and h[4] == s[4]  # window 0 at t = -(n-1) dt, 1 at t = 0
    assert np.array_equal(f, np.concatenate([h[4:], h[:4]]))


def test_w2t_inverts_t2w():
    rng = np.random.default_rng(6)
    x = rng.normal(size=30) + 1j * rng.normal(size=30)
    assert np.allclose(so.w2t(so.t2w(x, 0.2), 0.2), x, atol=1e-13)
    # t2w of a delta at j = 0 is dt at every frequency
    d = np.zeros(16, np.complex128)
    d[0] = 1
    assert np.allclose(so.t2w(d, 0.3), 0.3)


def test_read_signal_rows_and_dt(tmp_path):
    """Header lines skipped, extra columns ignored, rows read until t > MaxTime."""
    p = tmp_path / "corr.txt"
    with open(p, "w") as fh:
        fh.write("s Re(O1(1S0)) Im(O1(1S0)) Re(O1(3S1)) Im(O1(3S1))\n")
        for k in range(21):
            fh.write(f"{k * 0.1:+.9E} {np.cos(k):+.9E} {np.sin(k):+.9E} 0 0\n")
    sig, dt = so.read_signal(p, 1, 1.0)
    assert sig.size == 11 and dt == pytest.approx(0.1, abs=1e-12)
    assert np.allclose(sig, np.exp(1j * np.arange(11)), atol=1e-9)


def test_write_order():
    spec = np.arange(8.0)
    rows = so.write_order(spec, 0.5)
    dw = 2 * np.pi / (0.5 * 8)
    assert rows[0, 0] == pytest.approx(-np.pi / 0.5)
    assert np.allclose(np.diff(rows[:, 0]), dw)
    assert list(rows[:, 1]) == [4, 5, 6, 7, 0, 1, 2, 3]
