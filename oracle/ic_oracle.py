"""CPU restatement of the transverse-polarised initial conditions with a
general occupation (test infrastructure).

Only tests use this module, as the checker of ``nrqcd_amd/csrc/nrq_ic.hip``;
the product never imports it.  It assembles the fields in numpy from the C
oracle's pinned pieces (ranlxd stream, momenta, polarisation vectors, group
exponential, x2p/p2x).

Restated from the reference (``src/gaugeconfiguration_su3.f90``):

* ``TransversePolarisedOccupiedInit`` (``:1263-1446``): per lattice site (in
  lattice order) and polarisation, 8 complex normals of which the real parts
  are kept; those 16 real fields go to momentum space (x2p); at |p| > 1e-12
  A^a_i(p) = sqrt(2 f/|p|) sum_pol r^a_pol(p) eps_i,pol(p) and
  E^a_i(p) = i sqrt(2 f |p|) (same sum).
* ``TransversePolarisedOccupiedInit_specificSender`` (``:959-1255``): per
  lattice index with |p| > 1e-12, in lattice order, and per generator, 8
  complex normals (16 uniforms) are drawn and the first two are the
  polarisation amplitudes (complex, drawn directly in momentum space); the
  prefactors carry the physical volume, sqrt(V f / |p| / 2) and
  i sqrt(V f |p| / 2).  Then every component is made Hermitian in momentum
  space in lattice order (``SymmetriseInPspace``, ``:1163-1255``): the entry of
  the smaller of the indices p, -p is kept, the other becomes its conjugate,
  and a self-conjugate momentum keeps its real part.
* Both: the optional correlators (``:1390-1421``, ``:1105-1136``) are
  sum_a Re[sum_ij X_i P_ij conj(X_j)] / V / 8 of the momentum-space fields with
  the real transverse projector; then p2x, U_i = exp(i a_i Re A^a T^a),
  E^a_i = a0 a_i Re E^a.

Pinning (tests/test_oracle_ic.py): with the box occupation the first form
reproduces the C oracle's box initial condition (itself pinned against the
reference's compiled pieces and the lattice Coulomb condition); the second
gives real position-space fields before the real part is taken, satisfies
the lattice Coulomb condition, and a single occupied mode gives the
closed-form plane wave of the drawn amplitudes.
"""
from __future__ import annotations

import numpy as np

ZERO_TOL = 1e-12  # src/tolerances.f90


def box_occupation(pnorm, qs, amp, coupling):
    """GetBoxOccupation (:1455-1471)"""
    return np.where(pnorm < qs, amp, coupling * coupling / 2)


def _momenta(orc, vol):
    p = np.array([orc.momentum(k + 1) for k in range(vol)])
    pn = np.array([orc.norm2_momentum(k + 1) for k in range(vol)])
    return p, pn


def _real_projector(p, pn):
    P = np.zeros((len(pn), 3, 3))
    nz = pn > ZERO_TOL
    P[nz] = np.real(np.eye(3)[None] - np.conj(p[nz])[:, :, None] * p[nz][:, None, :] / (pn[nz] ** 2)[:, None, None]) / 2
    return P


def transverse_correlator(F, P, vol_phys):
    """F: (V, 3, 8) momentum-space field -> per index sum_a Re[F_i P_ij conj(F_j)] / V / 8"""
    return np.real(np.einsum("sia,sij,sja->s", F, P, np.conj(F))) / vol_phys / 8


def momentum_fields(orc, ext, spacing, occupation, specific_sender: bool):
    """(A, E) in momentum space, shape (V, 3, 8), after the symmetrisation of
    the specific-sender form."""
    vol = int(np.prod(ext))
    vol_phys = vol * spacing[1] * spacing[2] * spacing[3]
    p, pn = _momenta(orc, vol)
    A = np.zeros((vol, 3, 8), np.complex128)
    E = np.zeros((vol, 3, 8), np.complex128)
    if not specific_sender:
        u = orc.ranlxd(32 * vol).reshape(vol, 2, 16)
        r = np.sqrt(-2 * np.log(u[:, :, 0::2])) * np.cos(2 * np.pi * u[:, :, 1::2])  # (V, pol, a)
        rk = np.stack([np.stack([orc.x2p(r[:, pol, a]) for a in range(8)], 1) for pol in range(2)], 1)  # (V, pol, a)
    for s in range(vol):
        if not pn[s] > ZERO_TOL:
            continue
        eps = orc.polarisation(p[s])
        f = occupation[s]
        if specific_sender:
            u = orc.ranlxd(128).reshape(8, 16)  # per generator 8 complex normals, the first two used
            r = np.sqrt(-2 * np.log(u[:, [0, 2]])) * np.exp(2j * np.pi * u[:, [1, 3]])  # (a, pol)
            pa = np.sqrt(vol_phys * f / pn[s] / 2)
            pe = 1j * np.sqrt(vol_phys * f * pn[s] / 2)
        else:
            r = rk[s].T  # (a, pol)
            pa = np.sqrt(f / pn[s] * 2)
            pe = 1j * np.sqrt(f * pn[s] * 2)
        mode = eps[:, 0][:, None] * r[:, 0][None, :] + eps[:, 1][:, None] * r[:, 1][None, :]  # (i, a)
        A[s] = pa * mode
        E[s] = pe * mode
    if specific_sender:
        for s in range(vol):
            t = orc.negative_index(s + 1) - 1
            if t == s:
                A[s] = A[s].real
                E[s] = E[s].real
            elif t > s:
                A[t] = np.conj(A[s])
                E[t] = np.conj(E[s])
    return A, E


def transverse_init(orc, ext, spacing, occupation, specific_sender: bool):
    """(U (3,3,3,V), E (8,3,V) in the reference layout, aa, ee, A(x), E(x) complex)"""
    vol = int(np.prod(ext))
    vol_phys = vol * spacing[1] * spacing[2] * spacing[3]
    p, pn = _momenta(orc, vol)
    A, Ef = momentum_fields(orc, ext, spacing, occupation, specific_sender)
    P = _real_projector(p, pn)
    aa = transverse_correlator(A, P, vol_phys)
    ee = transverse_correlator(Ef, P, vol_phys)
    Ax = np.zeros_like(A)
    Ex = np.zeros_like(Ef)
    for i in range(3):
        for a in range(8):
            Ax[:, i, a] = orc.p2x(A[:, i, a])
            Ex[:, i, a] = orc.p2x(Ef[:, i, a])
    U = np.zeros((3, 3, 3, vol), np.complex128, order="F")
    Eo = np.zeros((8, 3, vol), order="F")
    for s in range(vol):
        for i in range(3):
            U[:, :, i, s] = orc.group_exp(Ax[s, i].real * spacing[i + 1])
            Eo[:, i, s] = Ex[s, i].real * spacing[0] * spacing[i + 1]
    return U, Eo, aa, ee, Ax, Ex
