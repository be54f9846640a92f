"""CPU restatement of the learned preconditioner's network (test infrastructure).

Only tests use this module, as the checker of the GPU network in
``nrqcd_amd/csrc/nrq_cnn.hip``; the product never imports it.

The network N maps 18 real channels per site to 18: four 3^3 periodic
convolutions (channels 18 -> 32 -> 32 -> 32 -> 18, padded to 32 everywhere),
ReLU after the first three, no biases.  Weights are [layer][tap][co][ci] with
tap = 9 (dz+1) + 3 (dy+1) + (dx+1); a field is an array [z][y][x][c] (site
index x + n0 (y + n1 z)).  With ``bf16=True`` the inputs, weights and hidden
activations are rounded to bfloat16 (round to nearest even) where the GPU
rounds them; sums are exact in float64 here (FP32 on the GPU).

Pinning: the analytic gradient is checked against central finite
differences of the loss (float64, no rounding) in tests/test_oracle_cnn.py;
the forward pass is a direct transcription of the convolution definition,
checked against an explicit per-site loop on a tiny lattice there.
"""
from __future__ import annotations

import numpy as np

CH, TAPS, NL, CIN = 32, 27, 4, 18
NW = NL * TAPS * CH * CH
OFFS = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]  # tap order


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _shift(x: np.ndarray, dx: int, dy: int, dz: int) -> np.ndarray:
    """y[z][y][x] = x[z+dz][y+dy][x+dx] (periodic)."""
    return np.roll(x, shift=(-dz, -dy, -dx), axis=(0, 1, 2))


def conv(x: np.ndarray, w: np.ndarray) -> np.ndarray:
    """out[s][co] = sum_{tap, ci} w[tap][co][ci] x[s + off(tap)][ci]."""
    out = np.zeros(x.shape[:3] + (w.shape[1],))
    for t, (dx, dy, dz) in enumerate(OFFS):
        out += _shift(x, dx, dy, dz) @ w[t].T
    return out


def conv_t(d: np.ndarray, w: np.ndarray) -> np.ndarray:
    """Adjoint of conv in its input: out[s][ci] = sum w[tap][co][ci] d[s - off(tap)][co]."""
    out = np.zeros(d.shape[:3] + (w.shape[2],))
    for t, (dx, dy, dz) in enumerate(OFFS):
        out += _shift(d, -dx, -dy, -dz) @ w[t]
    return out


def weights_layers(w: np.ndarray) -> list[np.ndarray]:
    return [np.asarray(w, dtype=np.float64).reshape(NL, TAPS, CH, CH)[l] for l in range(NL)]


def forward(inp: np.ndarray, w: np.ndarray, bf: bool = True):
    """inp [z][y][x][18] -> (output [z][y][x][18], layer inputs for backward)."""
    ws = weights_layers(w)
    if bf:
        ws = [bf16(x) for x in ws]
    a = np.zeros(inp.shape[:3] + (CH,))
    a[..., :CIN] = bf16(inp) if bf else inp
    acts = [a]
    for l in range(NL):
        o = conv(acts[-1], ws[l])
        if l < NL - 1:
            o = np.maximum(o, 0.0)
            acts.append(bf16(o) if bf else o)
        else:
            out = o
    return out[..., :CIN], acts, ws


def loss_grad(inp: np.ndarray, target: np.ndarray, w: np.ndarray, norm: float, bf: bool = True):
    """loss = sum (N(inp) - target)^2 / norm and d loss / d w (flat, NW)."""
    out, acts, ws = forward(inp, w, bf)
    diff = out - target
    loss = float(np.sum(diff * diff) / norm)
    d = np.zeros(out.shape[:3] + (CH,))
    d[..., :CIN] = 2.0 * diff / norm
    if bf:
        d = bf16(d)
    grad = np.zeros((NL, TAPS, CH, CH))
    for l in range(NL - 1, -1, -1):
        a = acts[l]
        for t, (dx, dy, dz) in enumerate(OFFS):
            sh = _shift(a, dx, dy, dz)
            grad[l, t] = np.einsum("zyxo,zyxi->oi", d, sh)
        if l > 0:
            dn = conv_t(d, ws[l])
            dn = np.where(acts[l] > 0, dn, 0.0)
            d = bf16(dn) if bf else dn
    return loss, grad.reshape(-1)


def random_weights(rng: np.random.Generator, scale: float = 1.0) -> np.ndarray:
    """He-scaled random weights with the padded channels zero."""
    w = np.zeros((NL, TAPS, CH, CH))
    for l in range(NL):
        cin = CIN if l == 0 else CH
        cout = CIN if l == NL - 1 else CH
        w[l, :, :cout, :cin] = rng.normal(size=(TAPS, cout, cin)) * np.sqrt(2.0 / (27 * cin)) * scale
    return w.reshape(-1)
