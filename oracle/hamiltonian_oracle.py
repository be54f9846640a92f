"""CPU restatement of the full NRQCD system matrix and Crank-Nicolson step
(test infrastructure).

Only tests use this module, as the checker of the general-Hamiltonian path
of ``nrqcd_amd/csrc/nrq_hamiltonian.hip``; the product never imports it.

``system_matrix`` assembles, block by block as a scipy sparse matrix,
BuildSystemMatrix (``src/nrqcd.f90:1175-1387``) on one rank:
S = 1 + (-i dt a0 / 2) H with dt the step width and

* c1 kinetic (``:1210-1242``): + c1/(M a_i^2) on the diagonal, hops
  -c1/(2 M a_i^2) C2CS(U_i(x)) to x + i and C2CS(U_i(x - i))^+ to x - i;
* c3 (``:1244-1270``): on the diagonal -c3/(8 M^2) C2CS(D_i(x)),
  D_i = (U_i(x) e_i(x + i) U_i(x)^+ - U_i(x - i)^+ e_i(x - i) U_i(x - i)) / (2 a_i),
  with e_i = -F_0i;
* c2 (``:1272-1282``): on the diagonal -c2/(2 M) C2CS(B_i(x)),
  B_i = (F_jk - F_kj)/2 (``:1669-1693``), no Pauli matrix (C2CS of a colour
  matrix, a reference quirk);
* c4 (``:1284-1381``): for (i, j, k) cyclic (epsilon = +1) and anticyclic
  (epsilon = -1), hops to x + i with epsilon (-c4)/(8 M^2) i
  Kron((e_j(x) U_i(x) + U_i(x) e_j(x + i)) / (2 a_i), sigma_k) and to x - i
  with epsilon (+c4)/(8 M^2) i Kron((e_j(x) W + W e_j(x - i)) / (2 a_i),
  sigma_k), W = U_i(x - i)^+.

The field strength is F_ij(x) = Re[ArgLogU(P_ij(x))] / (a_i a_j), the real
part taken element by element (``gaugeconfiguration_su3.f90:1563-1581``),
with the spatial plaquette P_ij or, for i = 0, the temporal plaquette
P_0j = exp(i E_j^a T^a) (``:1656-1679``).  ArgLogU (``matrixoperations.f90:
474-495``) is S^+ diag(atan2(Im, Re) of the eigenvalues) S.

``cn_update`` follows Update_CrankNicholson (``src/nrqcd.f90:1053-1168``):
per column, rhs = S psi and S^+ psi' = rhs, with S built for (M, c) for the
quark and (-M, conj(c)) for the antiquark; the solves are exact (sparse LU).

Pinning (tests/test_oracle_hamiltonian.py): with c = (c1, 0, 0, 0) the
matrix equals the C oracle's operator; ArgLogU against the reference's
compiled ArgLogU; the c2 term against a constant abelian magnetic field, the
c3 term against E fields without the group exponential/logarithm round trip,
the c4 term against its momentum-space multiplier on plane waves (U = 1,
constant E).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .gaugefix_oracle import Lattice, dag, links_sdrc


def arglogu(M):
    """ArgLogU of a batch of normal matrices (..., 3, 3)"""
    ev, S = np.linalg.eig(M)
    theta = np.arctan2(ev.imag, ev.real)
    return (S * theta[..., None, :]) @ np.linalg.inv(S)


def algebra(lat: Lattice, coords):
    """sum_a coords[..., a] T^a"""
    return np.einsum("...a,rca->...rc", coords, lat.T)


def exp_i(H):
    """exp(i H) for Hermitian H (batched)"""
    Hh = (H + dag(H)) / 2
    w, v = np.linalg.eigh(Hh)
    return (v * np.exp(1j * w)[..., None, :]) @ dag(v)


def field_strengths(lat: Lattice, W, E):
    """F0[j] (V,3,3) = F_0j and F[(i,j)] (V,3,3) for spatial i != j (0-based
    directions), with W the links [site, dir, row, col] and E (8, 3, V)."""
    Eal = np.transpose(np.asarray(E), (2, 1, 0))  # (V, dir, a)
    F0 = []
    for j in range(3):
        P = exp_i(algebra(lat, Eal[:, j, :]))
        F0.append(np.real(arglogu(P)) / (lat.a[0] * lat.a[j + 1]))
    F = {}
    for i in range(3):
        for j in range(3):
            if i == j:
                continue
            P = W[:, i] @ W[lat.up[i], j] @ dag(W[lat.up[j], i]) @ dag(W[:, j])
            F[(i, j)] = np.real(arglogu(P)) / (lat.a[i + 1] * lat.a[j + 1])
    return F0, F


def _c2cs(m):
    return np.kron(m, np.eye(2))


def system_matrix(lat: Lattice, U, E, mass, c, stepwidth, pauli):
    """S (6V x 6V, scipy csr), row 6 x + d with d = 2 colour + spin; pauli:
    the reference's three Pauli matrices (su2.f90:68-88, no 1/2)."""
    W = links_sdrc(U)
    V = lat.vol
    a0 = lat.a[0]
    pre = -1j * stepwidth / 2 * a0
    c = np.asarray(c, np.complex128)
    rows, cols, vals = [], [], []

    def add(x, y, block):
        r = 6 * x + np.arange(6)
        cc = 6 * y + np.arange(6)
        rr, cc2 = np.meshgrid(r, cc, indexing="ij")
        rows.append(rr.ravel())
        cols.append(cc2.ravel())
        vals.append(np.asarray(block, np.complex128).ravel())

    F0, F = field_strengths(lat, W, E) if np.any(c[1:] != 0) else (None, None)
    eye6 = np.eye(6)
    for x in range(V):
        add(x, x, eye6)
        for i in range(3):
            ai = lat.a[i + 1]
            xp, xm = lat.up[i][x], lat.dn[i][x]
            add(x, x, pre * c[0] / mass / ai ** 2 * eye6)
            add(x, xp, pre * (-c[0]) / 2 / mass / ai ** 2 * _c2cs(W[x, i]))
            add(x, xm, pre * (-c[0]) / 2 / mass / ai ** 2 * dag(_c2cs(W[xm, i])))
            if F0 is None:
                continue
            # c3
            ep, em = -F0[i][xp], -F0[i][xm]
            link, link_m = W[x, i], W[xm, i]
            D = (link @ ep @ dag(link) - dag(link_m) @ em @ link_m) / (2 * ai)
            add(x, x, pre * (-c[2]) / 8 / mass ** 2 * _c2cs(D))
            # c2
            j, k = (i + 1) % 3, (i + 2) % 3
            B = (F[(j, k)][x] - F[(k, j)][x]) / 2
            add(x, x, pre * (-c[1]) / 2 / mass * _c2cs(B))
            # c4
            for (j, k, lc) in (((i + 1) % 3, (i + 2) % 3, +1), ((i + 2) % 3, (i + 1) % 3, -1)):
                e, e_p, e_m = -F0[j][x], -F0[j][xp], -F0[j][xm]
                Dp = (e @ link + link @ e_p) / 2 / ai
                add(x, xp, pre * lc * (-c[3]) / 8 / mass ** 2 * 1j * np.kron(Dp, pauli[k]))
                wm = dag(W[xm, i])
                Dm = (e @ wm + wm @ e_m) / 2 / ai
                add(x, xm, pre * lc * (+c[3]) / 8 / mass ** 2 * 1j * np.kron(Dm, pauli[k]))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(6 * V, 6 * V))


def prop_to_vec(P, col):
    """column col of a propagator (6, 6, V) -> vector over 6 x + d"""
    return np.ascontiguousarray(np.asarray(P)[:, col, :].T).reshape(-1)


def vec_to_prop(P, col, v):
    P[:, col, :] = v.reshape(-1, 6).T


def cn_update(lat: Lattice, U, E, Q, A, mass, c, stepwidth, pauli):
    """One CN step of both propagators (in place), exact solves."""
    c = np.asarray(c, np.complex128)
    for P, m, cc in ((Q, mass, c), (A, -mass, np.conj(c))):
        S = system_matrix(lat, U, E, m, cc, stepwidth, pauli)
        lu = spla.splu(S.conj().T.tocsc())
        for col in range(6):
            v = prop_to_vec(P, col)
            vec_to_prop(P, col, lu.solve(S @ v))
