"""CPU restatement of the gluon-occupation measurement (test infrastructure).

Only tests use this module, as the checker of ``nrqcd_amd/csrc/nrq_gaugefix.hip``;
the product never imports it.

Restated from the reference:

* Coulomb gauge fixing, Davies' Fourier-accelerated steepest descent
  (``src/gaugeconfiguration_su3.f90:2269-2521``): per iteration the divergence
  Delta(x) = sum_i skew(U_i(x - i) - U_i(x)), skew(d) = (d - d^+)/2 (not made
  traceless, ``:2500-2521``); stop when max_x ||Delta(x)||_F / 9 <= tol
  (``:2417-2465``, ``src/matrixoperations.f90:47-52``); otherwise Delta is
  taken to momentum space (x2p), every mode with |p| > 1e-12 multiplied by
  (p_max / |p|)^2, taken back (p2x), multiplied by alpha, and g(x) =
  exp(Delta(x)) (ExpAH, ``src/matrixoperations.f90:547-567``) transforms the
  links only: U_i(x) <- g(x) U_i(x) g(x + i)^+ (``:2479-2497``).  E is not
  transformed, as in the reference.
* Transverse AA and EE correlators (``:1825-1972``): the fields
  A^a_i(x) = GetGroupLog(U_i(x))^a / a_i and E^a_i(x) / (a_i a0), each taken to
  momentum space with x2p, contracted with the transverse projector
  P_ij = Re[(delta_ij - conj(p_i) p_j / |p|^2) / 2] (0 at p = 0; the
  function is declared real, so only the real part of the complex expression
  survives, ``src/lattice.f90:1178-1195``) of the backward-derivative momentum
  (``src/lattice.f90:963-979``), divided by the physical volume and averaged
  over the 8 generators.  GetGroupLog (``src/su3.f90:318-334``) is
  2 Re Tr(T^a H) of H = S^+ diag(atan2(Im ev, Re ev)) S for the eigen-
  decomposition U = S^+ diag(ev) S (ArgLogU, ``src/matrixoperations.f90:474-495``).
* The gluon distribution written by the driver is sqrt(AA EE) per momentum,
  next to |p| (``src/main.f90:3093-3189``).

FFT conventions (``src/xpfft.f90:129-137``): x2p is the forward DFT (e^{-i p x})
scaled by a1 a2 a3, p2x the backward DFT scaled by 1 / (physical volume).

Pinning (tests/test_oracle_gaugefix.py): the group log against the reference's
compiled GetGroupLog (oracle/_ref), the exponential against the compiled ExpAH
on non-traceless skew-Hermitian input, x2p/p2x against the C oracle's DFT;
gauge fixing by invariants (a pure-gauge configuration is fixed to a constant,
the divergence falls below the tolerance, plaquettes are unchanged, the links
stay unitary); the correlators by a transverse plane wave with a closed-form
answer and by the vanishing of longitudinal modes.
"""
from __future__ import annotations

import numpy as np

ZERO_TOL = 1e-12  # src/tolerances.f90


class Lattice:
    """Extents (n0, n1, n2), spacings (a0, a1, a2, a3); site s = x + n0 (y + n1 z)."""

    def __init__(self, ext, spacing, generators):
        self.n = tuple(int(v) for v in ext)
        self.a = tuple(float(v) for v in spacing)
        self.vol = self.n[0] * self.n[1] * self.n[2]
        self.vol_phys = self.vol * self.a[1] * self.a[2] * self.a[3]
        self.T = np.asarray(generators)  # (3, 3, 8)
        n0, n1, n2 = self.n
        z, y, x = np.meshgrid(np.arange(n2), np.arange(n1), np.arange(n0), indexing="ij")
        pos = [x.ravel(), y.ravel(), z.ravel()]
        self.up, self.dn = [], []
        for d in range(3):
            q = [c.copy() for c in pos]
            q[d] = (q[d] + 1) % self.n[d]
            self.up.append(q[0] + n0 * (q[1] + n1 * q[2]))
            q = [c.copy() for c in pos]
            q[d] = (q[d] - 1) % self.n[d]
            self.dn.append(q[0] + n0 * (q[1] + n1 * q[2]))
        # backward-derivative momentum of every lattice site (as a momentum index)
        self.p = np.zeros((self.vol, 3), np.complex128)
        for d in range(3):
            arg = np.pi * pos[d] / self.n[d]
            self.p[:, d] = 2 * np.sin(arg) / self.a[d + 1] * np.exp(-1j * arg)
        self.pnorm = np.sqrt(np.sum(np.abs(self.p) ** 2, axis=1))
        self.pmax = float(self.pnorm.max())

    # FFTs of fields with the site index first: shape (V, ...)
    def _grid(self, f):
        n0, n1, n2 = self.n
        return f.reshape((n2, n1, n0) + f.shape[1:])

    def x2p(self, f):
        F = np.fft.fftn(self._grid(f.astype(np.complex128)), axes=(0, 1, 2))
        return F.reshape(f.shape) * (self.a[1] * self.a[2] * self.a[3])

    def p2x(self, f):
        F = np.fft.ifftn(self._grid(f.astype(np.complex128)), axes=(0, 1, 2)) * self.vol
        return F.reshape(f.shape) / self.vol_phys


def links_sdrc(U):
    """(3, 3, 3, V) Fortran links [row, col, dir, site] -> [site, dir, row, col]."""
    return np.ascontiguousarray(np.transpose(np.asarray(U), (3, 2, 0, 1)))


def links_fortran(W):
    return np.asfortranarray(np.transpose(W, (2, 3, 1, 0)))


def dag(m):
    return np.conj(np.swapaxes(m, -1, -2))


def divergence(lat: Lattice, W):
    """Delta(x) = sum_i skew(U_i(x - i) - U_i(x)), shape (V, 3, 3)."""
    out = np.zeros((lat.vol, 3, 3), np.complex128)
    for i in range(3):
        d = W[lat.dn[i], i] - W[:, i]
        out += (d - dag(d)) / 2
    return out


def deviation(Delta):
    """||Delta(x)||_F / 9 per site."""
    return np.sqrt(np.sum(np.abs(Delta) ** 2, axis=(1, 2))) / 9


def expah(A):
    """exp(A) for skew-Hermitian A (batched): exp(i H), H = -i A made Hermitian."""
    H = -1j * A
    H = (H + dag(H)) / 2
    w, v = np.linalg.eigh(H)
    return (v * np.exp(1j * w)[..., None, :]) @ dag(v)


def coulomb_gauge_fix(lat: Lattice, U, tol=1e-8, alpha=0.08, maxit=100):
    """Returns (links in Fortran layout, transformations applied, final max deviation)."""
    W = links_sdrc(U).astype(np.complex128)
    done = 0
    for _ in range(maxit):
        D = divergence(lat, W)
        dev = float(deviation(D).max())
        if dev <= tol:
            return links_fortran(W), done, dev
        F = lat.x2p(D)
        scale = np.ones(lat.vol)
        nz = lat.pnorm > ZERO_TOL
        scale[nz] = (lat.pmax / lat.pnorm[nz]) ** 2
        D = lat.p2x(F * scale[:, None, None]) * alpha
        g = expah(D)
        for i in range(3):
            W[:, i] = g @ W[:, i] @ dag(g[lat.up[i]])
        done += 1
    D = divergence(lat, W)
    return links_fortran(W), done, float(deviation(D).max())


def group_log_coords(lat: Lattice, M):
    """GetGroupLog of a batch of unitary matrices (N, 3, 3) -> (N, 8)."""
    ev, S = np.linalg.eig(M)  # M = S diag(ev) S^-1
    theta = np.arctan2(ev.imag, ev.real)
    H = (S * theta[..., None, :]) @ np.linalg.inv(S)
    # 2 Re Tr(T^a H) = 2 Re sum_jk T^a_jk H_kj
    return 2 * np.real(np.einsum("jka,nkj->na", lat.T, H))


def transverse_correlator(lat: Lattice, F):
    """F: real fields (V, 3 directions, 8 generators) -> correlator per momentum (V,)."""
    P = np.zeros((lat.vol, 3, 3), np.complex128)
    nz = lat.pnorm > ZERO_TOL
    p = lat.p[nz]
    P[nz] = (np.eye(3)[None] - np.conj(p)[:, :, None] * p[:, None, :] / (lat.pnorm[nz] ** 2)[:, None, None]) / 2
    P = np.real(P)
    Fp = lat.x2p(F)  # (V, 3, 8)
    res = np.einsum("sia,sij,sja->s", Fp, P, np.conj(Fp))
    return np.real(res) / lat.vol_phys / 8


def transverse_correlators(lat: Lattice, U, E):
    """(AA, EE) per momentum index for links U (3,3,3,V) and E (8,3,V)."""
    W = links_sdrc(U)
    A = np.zeros((lat.vol, 3, 8))
    for i in range(3):
        A[:, i, :] = group_log_coords(lat, W[:, i]) / lat.a[i + 1]
    Ef = np.transpose(np.asarray(E), (2, 1, 0)) / np.array(lat.a[1:4])[None, :, None] / lat.a[0]
    return transverse_correlator(lat, A), transverse_correlator(lat, Ef)


def plaquettes(lat: Lattice, U):
    """Re Tr P_ij(x) for all i < j, shape (V, 3)."""
    W = links_sdrc(U)
    out = []
    for i, j in ((0, 1), (0, 2), (1, 2)):
        P = W[:, i] @ W[lat.up[i], j] @ dag(W[lat.up[j], i]) @ dag(W[:, j])
        out.append(np.real(np.trace(P, axis1=1, axis2=2)))
    return np.stack(out, axis=1)
