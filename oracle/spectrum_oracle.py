"""CPU restatement of the spectral-function post-processing (test
infrastructure).

Only tests use this module, as the checker of ``nrqcd_amd/csrc/nrq_spectrum.hip``
and of the driver's mode 4; the product never imports it.

Restated from the reference (``src/main.f90:1865-2100``, ``src/windowing.f90``,
``src/twfft.f90``):

* ``read_signal``: ReadSignal (``main.f90:2018-2062``).  After ``nheader``
  header lines, rows ``t re im`` are read list-directed until a row with
  t > MaxTime or a read error; dt is the difference of the first two times.
* ``symmetrise``: Symmetrisation (``:1970-1992``).  n points become 2n - 2:
  entries 1 .. n-1 are -conj(C(t_{n-1})) .. -conj(C(t_1)), entries
  n .. 2n-2 are C(t_0) .. C(t_{n-2}).
* ``hann``: ApplyHannWindow (``windowing.f90:11-25``), x_i sin^2(pi (i-1) / N).
* ``symm2fft``: Symm2FFTformat (``main.f90:2000-2016``), the halves swapped.
* ``t2w``/``w2t``: the MKL DFTI backward transform (exp(+i), scale dt) and
  the forward transform (exp(-i), scale 1/(dt N)) (``twfft.f90:33-105``).
* ``spectrum``: 2 Im t2w(...) (``main.f90:1957-1961``).
* ``write_order``: WriteSpectrum2File (``main.f90:2066-2099``), negative
  frequencies -pi/dt + k dw first, then k dw, dw = 2 pi / (dt N).

Pinning (tests/test_oracle_spectrum.py): the spectrum of a single
exponential against its closed form as geometric series (independent of any
FFT), the real part of the transform (dt Re C(0): the symmetrised signal is
anti-Hermitian in time and the window vanishes at the unpaired point), w2t
inverting t2w, and the ReadSignal row selection on a file with a header and
rows past MaxTime.
"""
from __future__ import annotations

import numpy as np


def read_signal(path, nheader: int, maxtime: float):
    rows = []
    with open(path) as fh:
        lines = fh.read().splitlines()
    for line in lines[nheader:]:
        parts = line.replace(",", " ").split()
        if not parts:  # list-directed input skips empty records
            continue
        try:
            t, re, im = (float(v) for v in parts[:3])
        except ValueError:
            break
        if len(parts) < 3:
            break
        if t > maxtime:
            break
        rows.append((t, re, im))
    sig = np.array([complex(r[1], r[2]) for r in rows])
    dt = rows[1][0] - rows[0][0] if len(rows) > 1 else None
    return sig, dt


def symmetrise(c):
    c = np.asarray(c, np.complex128)
    n = c.size
    out = np.empty(2 * n - 2, np.complex128)
    out[: n - 1] = -np.conj(c[1:][::-1])
    out[n - 1:] = c[: n - 1]
    return out


def hann(x):
    N = x.size
    return x * np.sin(np.pi * np.arange(N) / N) ** 2


def symm2fft(x):
    N = x.size
    return np.concatenate([x[N // 2:], x[: N // 2]])


def t2w(x, dt):
    x = np.asarray(x, np.complex128)
    return np.fft.ifft(x) * x.size * dt


def w2t(x, dt):
    x = np.asarray(x, np.complex128)
    return np.fft.fft(x) / (dt * x.size)


def spectrum(c, dt):
    """(spectrum in transform order, (symmetrised, windowed, reordered))"""
    s = symmetrise(c)
    h = hann(s)
    f = symm2fft(h)
    return 2 * np.imag(t2w(f, dt)), (s, h, f)


def write_order(spec, dt):
    """(omega, rho) rows in the order WriteSpectrum2File writes them."""
    N = spec.size
    wmin, dw = -np.pi / dt, 2 * np.pi / (dt * N)
    rows = [(wmin + k * dw, spec[N // 2 + k]) for k in range(N - N // 2)]
    rows += [(k * dw, spec[k]) for k in range(N // 2)]
    return np.array(rows)
