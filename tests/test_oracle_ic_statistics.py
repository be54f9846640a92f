"""The oracle's box initial condition against the statistics the mathematics
fixes, and its anisotropic leapfrog and energy against closed forms.

Box initial condition (TransversePolarisedOccupiedInit,
gaugeconfiguration_su3.f90:1263-1446 with GetBoxOccupation :1455-1471).  Per
site, generator a and polarisation l the reference draws a real N(0, 1) noise
r_l^a(x) (real part of the Box-Muller normal, random.f90:238-253) and forms
  R_l(k) = D3 sum_x r_l(x) e^{-i k x}                  (x2p, xpfft.f90:134)
  A_i(k) = sqrt(2 f/|p|) sum_l R_l(k) eps_l,i(k)
  E_i(k) = i sqrt(2 f |p|) sum_l R_l(k) eps_l,i(k)     (:1353-1381)
  F(x)   = 1/(N D3) sum_k F(k) e^{+i k x}              (p2x, xpfft.f90:136)
and keeps alpha_i = a_i Re A_i(x), E_i = a0 a_i Re E_i(x) (:1431-1443), with
D3 = a1 a2 a3, N = N1 N2 N3, p(k) the backward-derivative momentum
p_j = 2 sin(pi n_j/N_j)/a_j e^{-i pi n_j/N_j} (lattice.f90:963-979).

The stored fields, transformed back with the same forward convention, are
  A~_i(k) = D3 sum_x (alpha_i(x)/a_i) e^{-ikx}
          = sqrt(2f/|p|)/2 sum_l R_l(k) [eps_l,i(k) + conj(eps_l,i(-k))]
and the same for E~ with i sqrt(2f|p|) and a minus sign.  Because
p(-k) = -conj p(k), the polarisation vectors of lattice.f90:1120-1176 obey
eps_1(-k) = -conj eps_1(k) and eps_2(-k) = conj eps_2(k): only eps_2
survives in A~ and only eps_1 in E~, so
  A~(k) = sqrt(2f/|p|) R_2(k) eps_2(k),   E~(k) = i sqrt(2f|p|) R_1(k) eps_1(k),
  <sum_i |A~^a_i(k)|^2> = (2f/|p|) D3^2 N,   <sum_i |E~^a_i(k)|^2> = 2f|p| D3^2 N
(E|R_l(k)|^2 = D3^2 N), A~ has no eps_1 part and E~ no eps_2 part, and the
eps_2 part of A~ and the eps_1 part of E~ are uncorrelated (independent noise).
The polarisation vectors used here are the cross-product forms
eps_1 = z x p/|p12| (or -x x p/|p23| when p_1 = 0), eps_2 = conj(eps_1) x p^,
which test_oracle_math pins to lattice.f90:1120-1176.  Each test plants a
plausible mistake into the oracle and shows the check rejects it.
"""
import numpy as np
import pytest

# f = 0.5 inside the box, g^2/2 = 0.125 outside; Qs near the median |p| of
# the lattices below, so that both sides hold many modes
BOX = dict(qs=2.0, amp=0.5, g=0.5)


def _momenta(ext, a):
    """Backward-derivative momentum of every lattice index (x fastest)."""
    n = np.array(ext)
    pos = np.stack(np.meshgrid(*[np.arange(e) for e in ext], indexing="ij"), -1)
    pos = pos.transpose(2, 1, 0, 3).reshape(-1, 3)
    return 2 * np.sin(np.pi * pos / n) / a * np.exp(-1j * np.pi * pos / n)


def _polarisations(p):
    pn = np.sqrt(np.sum(np.abs(p) ** 2, 1))
    on1 = np.abs(p[:, 0]) > 1e-12
    e1 = np.empty_like(p)
    p12 = np.sqrt(np.abs(p[:, 0]) ** 2 + np.abs(p[:, 1]) ** 2)
    p23 = np.sqrt(np.abs(p[:, 1]) ** 2 + np.abs(p[:, 2]) ** 2)
    e1[on1] = np.cross([0.0, 0.0, 1.0], p[on1]) / p12[on1][:, None]
    e1[~on1] = -np.cross([1.0, 0.0, 0.0], p[~on1]) / p23[~on1][:, None]
    e2 = np.cross(np.conj(e1), p / pn[:, None])
    return e1, e2


def _algebra_of_links(U, T):
    """alpha with U = exp(i alpha^a T^a), principal eigenphases (numpy eig);
    returns alpha (V, 3, 8) and the largest |eigenphase|."""
    V = U.shape[3]
    M = U.transpose(3, 2, 0, 1).reshape(-1, 3, 3)
    w, vec = np.linalg.eig(M)
    th = np.angle(w)
    H = np.einsum("nij,nj,njk->nik", vec, th, np.linalg.inv(vec))
    al = 2 * np.real(np.einsum("nij,jia->na", H, T))
    return al.reshape(V, 3, 8), float(np.max(np.abs(th)))


def ic_statistics(oracle, ext, sp, seeds, qs, amp, g):
    """Measured spectra of oracle.box_init over `seeds`, normalised by the
    closed forms of the module docstring."""
    oracle.lattice_init(ext, sp)
    V = oracle.vol
    a0, a = sp[0], np.array(sp[1:])
    D3 = float(np.prod(a))
    T = oracle.generators()
    p = _momenta(ext, a)
    pn = np.sqrt(np.sum(np.abs(p) ** 2, 1))
    nz = pn > 1e-12
    f = np.where(pn < qs, amp, g * g / 2)[nz]
    e1, e2 = _polarisations(p[nz])

    def fwd(field):
        return np.fft.fftn(field.reshape(ext[::-1])).reshape(V)[nz] * D3

    ra, re, perp_a, perp_e, cross, na, ne, thmax = [], [], 0.0, 0.0, 0j, 0.0, 0.0, 0.0
    for seed in seeds:
        U, E = oracle.box_init(seed, qs, amp, g)
        al, th = _algebra_of_links(U, T)
        thmax = max(thmax, th)
        At = np.array([[fwd(al[:, i, c] / a[i]) for c in range(8)] for i in range(3)])
        Et = np.array([[fwd(E[c, i, :] / (a0 * a[i])) for c in range(8)] for i in range(3)])
        ra.append(np.sum(np.abs(At) ** 2, 0) / (2 * f / pn[nz] * D3 ** 2 * V))
        re.append(np.sum(np.abs(Et) ** 2, 0) / (2 * f * pn[nz] * D3 ** 2 * V))
        a1 = np.einsum("ick,ki->ck", At, np.conj(e1))
        a2 = np.einsum("ick,ki->ck", At, np.conj(e2))
        x1 = np.einsum("ick,ki->ck", Et, np.conj(e1))
        x2 = np.einsum("ick,ki->ck", Et, np.conj(e2))
        perp_a = max(perp_a, np.max(np.abs(a1)) / np.max(np.abs(a2)))
        perp_e = max(perp_e, np.max(np.abs(x2)) / np.max(np.abs(x1)))
        cross += np.sum(x1 * np.conj(a2))
        na += np.sum(np.abs(a2) ** 2)
        ne += np.sum(np.abs(x1) ** 2)
    ra, re = np.stack(ra), np.stack(re)  # (seed, colour, mode)
    edges = np.quantile(pn[nz], np.linspace(0, 1, 7)[1:-1])
    shell = np.digitize(pn[nz], edges)
    inside = pn[nz] < qs
    groups = [shell == s for s in range(6)] + [inside, ~inside]
    # each normalised |R(k)|^2 has variance 1 (2 for the real, self-conjugate
    # momenta), so a group of n samples has a standard error <= sqrt(2/n)
    return {
        "A": float(ra.mean()), "E": float(re.mean()),
        "A_groups": [float(ra[:, :, m].mean()) for m in groups],
        "E_groups": [float(re[:, :, m].mean()) for m in groups],
        "group_err": [float(np.sqrt(2.0 / (ra.shape[0] * ra.shape[1] * m.sum()))) for m in groups],
        "perp_A": perp_a, "perp_E": perp_e,
        "corr": float(abs(cross) / np.sqrt(na * ne)),
        "max_eigenphase": thmax,
    }


def ic_ok(st, tol_mean=0.03, nsigma=5.0):
    """The closed forms hold: normalised spectra 1 overall and per |p| shell
    (and inside / outside the box, within nsigma standard errors), A without
    eps_1, E without eps_2, and uncorrelated A and E noise."""
    groups_ok = all(abs(v - 1) < nsigma * err
                    for vals in (st["A_groups"], st["E_groups"]) for v, err in zip(vals, st["group_err"]))
    return (abs(st["A"] - 1) < tol_mean and abs(st["E"] - 1) < tol_mean and groups_ok
            and st["perp_A"] < 1e-10 and st["perp_E"] < 1e-10 and st["corr"] < 0.05)


CASES = [((8, 8, 8), (0.1, 0.7, 0.9, 1.2)), ((6, 8, 10), (0.05, 1.0, 1.0, 1.0))]


@pytest.mark.parametrize("ext,sp", CASES)
def test_box_ic_spectra_match_closed_forms(oracle, ext, sp):
    st = ic_statistics(oracle, ext, sp, range(1, 17), **BOX)
    # the principal logarithm is the algebra element only below pi
    assert st["max_eigenphase"] < 2.5
    assert ic_ok(st), st


@pytest.mark.parametrize("plant,what", [(3, "sqrt(2) amplitude error"), (4, "|p| powers swapped"),
                                        (5, "polarisation noise swapped in E"), (6, "factor i dropped in E")])
def test_box_ic_statistics_reject_planted_mistakes(oracle, plant, what):
    ext, sp = CASES[0]
    oracle.set_plant(plant)
    try:
        st = ic_statistics(oracle, ext, sp, range(1, 9), **BOX)
    finally:
        oracle.set_plant(0)
    assert not ic_ok(st), f"planted mistake not detected: {what}: {st}"


# --------------------------------------------------------------------------
# anisotropic spacings: force and energy factors
# --------------------------------------------------------------------------
ANISO = (0.2, 1.0, 0.8, 1.3)


def _shift(Ui, j, ext):
    """U(x + j) for a (3, 3, V) field (x fastest)."""
    V = Ui.shape[-1]
    return np.roll(Ui.reshape(3, 3, *ext[::-1]), -1, axis=4 - j).reshape(3, 3, V)


def physical_hamiltonian(U, E, ext, sp):
    """sum_x [sum_i (E_i/(a0 a_i))^2/2 + sum_{i<j} 2 N_c (1 - Re Tr P_ij/N_c)/(a_i a_j)^2]:
    the classical Yang-Mills Hamiltonian per unit cell volume in the
    reference's lattice variables (link angle a_i A_i, E-field a0 a_i E_i,
    gaugeconfiguration_su3.f90:27-28), with the plaquettes in numpy."""
    a0, a = sp[0], sp[1:]
    h = sum(np.sum((E[:, i, :] / (a[i] * a0)) ** 2) / 2 for i in range(3))
    for i in range(3):
        for j in range(i + 1, 3):
            Ui, Uj = U[:, :, i, :], U[:, :, j, :]
            P = np.einsum("abx,bcx,dcx,edx->aex", Ui, _shift(Uj, i, ext), np.conj(_shift(Ui, j, ext)), np.conj(Uj))
            tr = np.real(np.einsum("aax->x", P))
            h += np.sum(6 * (1 - tr / 3)) / (a[i] * a[j]) ** 2
    return h


def _hamiltonian_deviation(oracle, ext, sp):
    """Largest relative deviation of the synchronised physical Hamiltonian
    over the same physical time at StepWidth 0.5 and 0.25."""
    oracle.lattice_init(ext, sp)
    U0, E0 = oracle.box_init(1, 0.5, 1.0, 1.0)
    devs = []
    for sw, n in ((0.5, 20), (0.25, 40)):
        U, E = U0.copy(order="F"), E0.copy(order="F")
        oracle.update_links(U, E, sw)
        ep = E.copy(order="F")
        oracle.update_efield(U, E, sw)
        h0 = physical_hamiltonian(U, (ep + E) / 2, ext, sp)
        d = 0.0
        for _ in range(n):
            oracle.update_links(U, E, sw)
            ep = E.copy(order="F")
            oracle.update_efield(U, E, sw)
            d = max(d, abs(physical_hamiltonian(U, (ep + E) / 2, ext, sp) - h0) / h0)
        devs.append(d)
    return devs


def test_anisotropic_leapfrog_conserves_the_hamiltonian(oracle):
    """The E update's staple coefficient StepWidth (a0/a_k)^2
    (gaugeconfiguration_su3.f90:2180-2186) is what Hamilton's equations give
    for the physical Hamiltonian with link angles a_i A_i: on a lattice with
    three different spatial spacings the synchronised Hamiltonian is
    conserved at O(dt^2) (halving StepWidth quarters the deviation)."""
    d = _hamiltonian_deviation(oracle, (4, 5, 6), ANISO)
    assert d[0] < 0.01 and 3.0 < d[0] / d[1] < 5.0, d


def test_anisotropic_staple_coefficient_planted_mistake_is_detected(oracle):
    oracle.set_plant(1)  # (a0/a_i)^2 in place of (a0/a_k)^2
    try:
        d = _hamiltonian_deviation(oracle, (4, 5, 6), ANISO)
    finally:
        oracle.set_plant(0)
    assert not (d[0] < 0.01 and 3.0 < d[0] / d[1] < 5.0), d


def test_reference_energy_is_not_the_conserved_quantity_on_anisotropic_lattices(oracle):
    """GetEnergy (gaugeconfiguration_su3.f90:1769-1817) divides the
    magnetic term by a_i a_j, not (a_i a_j)^2: on an anisotropic lattice the
    reference's energy drifts at O(1) while the physical Hamiltonian above is
    conserved.  The build keeps the reference's formula; this records the
    difference (both agree when every a_i = 1)."""
    oracle.lattice_init((4, 5, 6), ANISO)
    U, E = oracle.box_init(1, 0.5, 1.0, 1.0)
    e0, h0 = oracle.energy(U, E), physical_hamiltonian(U, E, (4, 5, 6), ANISO)
    oracle.leapfrog(U, E, 0.25, 40)
    e1, h1 = oracle.energy(U, E), physical_hamiltonian(U, E, (4, 5, 6), ANISO)
    assert abs(e1 - e0) / e0 > 10 * abs(h1 - h0) / h0
    oracle.lattice_init((4, 5, 6), (0.2, 1.0, 1.0, 1.0))
    U, E = oracle.box_init(1, 0.5, 1.0, 1.0)
    assert oracle.energy(U, E) == pytest.approx(physical_hamiltonian(U, E, (4, 5, 6), (0.2, 1.0, 1.0, 1.0)), rel=1e-13)


def _energy_closed_form_case(oracle, sp, theta, eps):
    """Cold lattice with U_1(x0) = exp(i theta T^3) = diag(e^{i th/2}, e^{-i th/2}, 1)
    and E^5_2(x1) = eps: four plaquettes touch the link, two in the (1,2)
    and two in the (1,3) plane, each with Re Tr P = 2 cos(th/2) + 1, so
      GetEnergy = eps^2 / (2 a2^2 a0^2) + 4 (1 - cos(th/2)) (2/(a1 a2) + 2/(a1 a3))."""
    ext = (4, 5, 6)
    oracle.lattice_init(ext, sp)
    U, E = oracle.cold()
    x0, x1 = 37, 71
    U[:, :, 0, x0 - 1] = np.diag([np.exp(0.5j * theta), np.exp(-0.5j * theta), 1.0])
    E[4, 1, x1 - 1] = eps
    a0, a1, a2, a3 = sp
    expect = eps ** 2 / (2 * a2 ** 2 * a0 ** 2) + 4 * (1 - np.cos(theta / 2)) * (2 / (a1 * a2) + 2 / (a1 * a3))
    return oracle.energy(U, E), expect


def test_energy_spacing_factors_closed_form(oracle):
    got, expect = _energy_closed_form_case(oracle, ANISO, 0.7, 0.3)
    assert got == pytest.approx(expect, rel=1e-13)
    oracle.set_plant(2)  # magnetic term over a_i a_i
    try:
        bad, _ = _energy_closed_form_case(oracle, ANISO, 0.7, 0.3)
    finally:
        oracle.set_plant(0)
    assert abs(bad - expect) > 1e-3 * expect
