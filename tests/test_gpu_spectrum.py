"""GPU parity of the spectral-function post-processing (nrq_spectrum.hip)
against the CPU restatement oracle/spectrum_oracle.py.

Bar: spectra within 1e-12 of their largest magnitude (hipFFT against numpy's
FFT, both FP64); the symmetrised and reordered signals exact, the windowed
one within 1e-15 relative (the window's sine may differ in the last bit)."""
import numpy as np
import pytest

from oracle import spectrum_oracle as so

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


# n = 2: the shortest signal; 129: N = 256; 1000: N = 1998 = 2 3^3 37;
# 4116: N = 8230 = 2 5 823 (a large prime factor)
@pytest.mark.parametrize("n, batch", [(2, 1), (129, 5), (1000, 3), (4116, 2)])
def test_spectra_match_oracle(nrq, n, batch):
    _say(f"spectral functions n={n} batch={batch}")
    rng = np.random.default_rng(n)
    sig = rng.normal(size=(batch, n)) + 1j * rng.normal(size=(batch, n))
    dt = 0.05
    spec, st = nrq.spectral_functions(sig, dt, stages=True)
    assert spec.shape == (batch, 2 * n - 2) and st.shape == (batch, 3, 2 * n - 2)
    for b in range(batch):
        want, (s, h, f) = so.spectrum(sig[b], dt)
        assert np.max(np.abs(spec[b] - want)) <= 1e-12 * np.max(np.abs(want))
        assert np.array_equal(st[b, 0], s)
        assert np.allclose(st[b, 1], h, rtol=1e-15, atol=1e-300)
        assert np.array_equal(st[b, 2], st[b, 1][(np.arange(2 * n - 2) + n - 1) % (2 * n - 2)])
    # without stages, and one signal at a time: the same numbers
    assert np.array_equal(nrq.spectral_functions(sig, dt), spec)
    assert np.array_equal(nrq.spectral_functions(sig[-1], dt), spec[-1])


def test_single_exponential_peak(nrq):
    """C(t) = i exp(-i E t), the form of a meson correlator (C(0) = i):
    the symmetrised signal is i exp(-i E t) at all times, so with E on the
    frequency grid (bin 37 of N = 4096) the spectrum at the peak is
    2 dt sum_m cos^2(pi m / N) = N dt."""
    n, dt = 2049, 0.05
    N = 2 * n - 2
    E = 2.0 * np.pi * 37 / (dt * N)
    c = 1j * np.exp(-1j * E * np.arange(n) * dt)
    spec = nrq.spectral_functions(c, dt)
    k = int(np.argmax(spec))
    assert k == 37
    assert spec[k] == pytest.approx(N * dt, rel=1e-9)


def test_t2w_w2t(nrq):
    rng = np.random.default_rng(7)
    for n in (1, 17, 4096):
        x = rng.normal(size=n) + 1j * rng.normal(size=n)
        F = nrq.t2w(x, 0.2)
        assert np.allclose(F, so.t2w(x, 0.2), rtol=0, atol=1e-12 * np.max(np.abs(F)))
        assert np.allclose(nrq.w2t(F, 0.2), x, rtol=0, atol=1e-12 * np.max(np.abs(x)))


def test_rejects_bad_arguments(nrq):
    with pytest.raises(nrq.NRQError):
        nrq.spectral_functions(np.ones(1, np.complex128), 0.1)
    with pytest.raises(nrq.NRQError):
        nrq.spectral_functions(np.ones(4, np.complex128), 0.0)
    with pytest.raises(nrq.NRQError):
        nrq.w2t(np.ones(4, np.complex128), 0.0)
