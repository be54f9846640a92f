"""CPU restatement of EquilibriumInit (test infrastructure).

Only tests use this module, as the checker of
``nrqcd_amd/csrc/nrq_equilibrium.hip``; the product never imports it.

Restated from ``src/gaugeconfiguration_su3.f90:380-814`` (one rank, no
measurement branches):

* cold links; sigma = sqrt(4 / beta); kappa = 0.01, kappa_dt = kappa a0;
* per outer iteration (nefieldinit of them): E = 0, then per site in lattice
  order and direction k: 8 complex normals (16 uniforms, Box-Muller), and
  E^a_k(x) = Re(sigma r^a) a0 (``:455-466``);
* G(x) = [sum_i (E_i(x) - U_i(x-i)^+ E_i(x-i) U_i(x-i)) / a_i - (charge term)] / a0
  with E as algebra matrices (GetG, ``:781-813``; the charge term is
  sum_a T^a Tr(T^a diag(rho(x)) / (a1 a2 a3))), and the deviation
  max_{x,a} 2 |Tr(T^a G(x))| (GetGdeviation, ``:724-779``);
* while the deviation exceeds the tolerance (default 1E-3, a single-precision
  literal: 0.0010000000474974513): from the previous configuration,
  E_k(x) <- coordinates of kappa_dt (U_k(x) G(x) U_k(x)^+ G(x+k) - G(x)) + E_k(x)
  (``:480-507``; a product of G(x) and G(x+k), taken literally);
* then nequilibrium leapfrog steps of step width 1 (``:696-698``).

Pinning (tests/test_oracle_equilibrium.py): G's deviation equals the C
oracle's Gauss measure (GetGdeviation, itself pinned); a configuration that
satisfies the Gauss law is a fixed point of the projection; the projection
sweep agrees with its adjoint-representation form (R(U) g and the d-symbols
of SU(3), neighbour from the site coordinates), which a dropped -G term, a
flipped sign, G(x-k), U^+ G U or kappa without a0 each fail; the E draws are
the Box-Muller normals of the pinned ranlxd stream at the predicted offsets.
"""
from __future__ import annotations

import numpy as np

from .gaugefix_oracle import Lattice, dag, links_sdrc

KAPPA = 0.01
DEFAULT_TOL = float(np.float32(1e-3))


def algebra(lat: Lattice, coords):
    return np.einsum("...a,rca->...rc", coords, lat.T)


def coordinates(lat: Lattice, M):
    """2 Re Tr(T^a M) (GetAlgebraCoordinate) -> (..., 8)"""
    return 2 * np.real(np.einsum("rca,...cr->...a", lat.T, M))


def gauss_matrices(lat: Lattice, W, E, rho=None):
    """G(x) (V, 3, 3); W links [site, dir, row, col], E (8, 3, V)"""
    Eal = algebra(lat, np.transpose(np.asarray(E), (2, 1, 0)))  # (V, dir, 3, 3)
    G = np.zeros((lat.vol, 3, 3), np.complex128)
    for i in range(3):
        m = lat.dn[i]
        G += (Eal[:, i] - dag(W[m, i]) @ Eal[m, i] @ W[m, i]) / lat.a[i + 1]
    if rho is not None:
        C = np.zeros((lat.vol, 3, 3), np.complex128)
        for c in range(3):
            C[:, c, c] = np.asarray(rho)[c] / (lat.a[1] * lat.a[2] * lat.a[3])
        tr = np.einsum("rca,xcr->xa", lat.T, C)
        G -= np.einsum("xa,rca->xrc", tr, lat.T)
    return G / lat.a[0]


def deviation(lat: Lattice, G):
    tr = np.einsum("rca,xcr->xa", lat.T, G)
    return float(np.max(2 * np.abs(tr)))


def project(lat: Lattice, W, E, G):
    """One projection sweep from the previous configuration (E (8,3,V))."""
    Eal = algebra(lat, np.transpose(np.asarray(E), (2, 1, 0)))
    kdt = KAPPA * lat.a[0]
    out = np.zeros_like(np.asarray(E))
    for k in range(3):
        U = W[:, k]
        term = kdt * (U @ G @ dag(U) @ G[lat.up[k]] - G) + Eal[:, k]
        out[:, k, :] = coordinates(lat, term).T
    return np.asfortranarray(out)


def equilibrium_init(orc, lat: Lattice, beta, nefieldinit, nequilibrium, tol=None, rho=None, max_projections=1000000):
    """Returns (U, E, records) with records[it] = (projections, energy, deviation after the leapfrog)."""
    tol = DEFAULT_TOL if tol is None else tol
    U, E = orc.cold()
    sigma = np.sqrt(4 / beta)
    a0 = lat.a[0]
    records = []
    for _ in range(nefieldinit):
        E[:] = 0
        u = orc.ranlxd(48 * lat.vol).reshape(lat.vol, 3, 16)
        re = np.sqrt(-2 * np.log(u[:, :, 0::2])) * np.cos(2 * np.pi * u[:, :, 1::2])  # (V, k, a)
        E[:] = np.transpose((sigma * re) * a0, (2, 1, 0))
        W = links_sdrc(U)
        G = gauss_matrices(lat, W, E, rho)
        dev = deviation(lat, G)
        nproj = 0
        while dev > tol:
            nproj += 1
            if nproj > max_projections:
                raise RuntimeError("projection did not converge")
            E[:] = project(lat, W, E, G)
            G = gauss_matrices(lat, W, E, rho)
            dev = deviation(lat, G)
        orc.leapfrog(U, E, 1.0, nequilibrium)
        G = gauss_matrices(lat, links_sdrc(U), E, rho)
        records.append((nproj, orc.energy(U, E), deviation(lat, G)))
    return U, E, records
