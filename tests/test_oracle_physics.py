"""The oracle's leapfrog, observables and Crank-Nicolson step against
invariants and closed forms (README.md:4-15 lists the reference's own checks:
energy conservation, Gauss law, norm conservation, forward-backward
reversibility, free spectral function)."""
import numpy as np
import pytest
import scipy.linalg

SP = (0.05, 1.0, 1.0, 1.0)


def _random_su3(rng, n):
    out = np.empty((n, 3, 3), complex)
    for k in range(n):
        z = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        q, r = np.linalg.qr(z)
        q = q * (np.diag(r) / np.abs(np.diag(r)))
        out[k] = q / np.linalg.det(q) ** (1 / 3)
    return out


def _gauge_transform(oracle, U, E, G):
    """U_i(x) -> g(x) U_i(x) g(x+i)^+, E_i(x) -> g(x) E_i(x) g(x)^+ (E as an
    algebra matrix)."""
    T = oracle.generators()
    U2, E2 = U.copy(order="F"), E.copy(order="F")
    for x in range(oracle.vol):
        for i in range(3):
            xp = oracle.neib(i + 1, x + 1) - 1
            U2[:, :, i, x] = G[x] @ U[:, :, i, x] @ G[xp].conj().T
            M = np.einsum("a,rca->rc", E[:, i, x], T)
            M2 = G[x] @ M @ G[x].conj().T
            E2[:, i, x] = [2 * np.real(np.trace(T[:, :, a] @ M2)) for a in range(8)]
    return U2, E2


def test_cold_configuration_is_a_fixed_point(oracle):
    oracle.lattice_init((4, 4, 4), SP)
    U, E = oracle.cold()
    oracle.leapfrog(U, E, 1.0, 3)
    assert np.array_equal(U, oracle.cold()[0]) and not E.any()
    assert oracle.energy(U, E) == 0.0


def test_single_link_excitation_by_hand(oracle):
    """Cold lattice with E^3_1(x0) = eps: one step gives U_1(x0) =
    diag(e^{i th/2}, e^{-i th/2}, 1), th = eps, and then
    E^3_1(x0) = eps - 8 a0^2 sin(th/2) and E^3_2(x0) = +2 a0^2 sin(th/2)
    (staple structure of gaugeconfiguration_su3.f90:2164-2258 with unit links)."""
    oracle.lattice_init((4, 4, 4), SP)
    U, E = oracle.cold()
    eps, x0 = 0.3, 22
    E[2, 0, x0 - 1] = eps
    oracle.leapfrog(U, E, 1.0, 1)
    th = eps
    assert np.allclose(U[:, :, 0, x0 - 1], np.diag([np.exp(0.5j * th), np.exp(-0.5j * th), 1]), atol=1e-15)
    a0 = SP[0]
    assert E[2, 0, x0 - 1] == pytest.approx(eps - 8 * a0 ** 2 * np.sin(th / 2), abs=1e-15)
    assert E[2, 1, x0 - 1] == pytest.approx(+2 * a0 ** 2 * np.sin(th / 2), abs=1e-15)
    assert E[2, 2, x0 - 1] == pytest.approx(+2 * a0 ** 2 * np.sin(th / 2), abs=1e-15)
    # x0 - 2 (the link U_2(x0 - 2) has the excited link as U_1(x+2)^+ ... staple)
    for a in (0, 1, 3, 4, 5, 6, 7):
        assert abs(E[a, 0, x0 - 1]) < 1e-16


def test_time_reversibility(oracle):
    oracle.lattice_init((6, 6, 6), SP)
    U, E = oracle.box_init(1, 0.5, 1.0, 1.0)
    U0, E0 = U.copy(order="F"), E.copy(order="F")
    oracle.leapfrog(U, E, 1.0, 60)
    oracle.leapfrog(U, E, -1.0, 60)
    assert np.max(np.abs(U - U0)) < 1e-12
    assert np.max(np.abs(E - E0)) < 1e-12


def test_unitarity_preserved(oracle):
    oracle.lattice_init((4, 4, 4), SP)
    U, E = oracle.box_init(2, 0.5, 1.0, 1.0)
    oracle.leapfrog(U, E, 1.0, 40)
    for x in range(oracle.vol):
        for i in range(3):
            u = U[:, :, i, x]
            assert np.max(np.abs(u @ u.conj().T - np.eye(3))) < 1e-13
            assert abs(np.linalg.det(u) - 1) < 1e-13


def test_gauge_covariance_of_leapfrog(oracle):
    oracle.lattice_init((4, 3, 5), SP)
    rng = np.random.default_rng(4)
    U, E = oracle.box_init(3, 0.5, 1.0, 1.0)
    G = _random_su3(rng, oracle.vol)
    Ug, Eg = _gauge_transform(oracle, U, E, G)
    oracle.leapfrog(U, E, 1.0, 3)
    oracle.leapfrog(Ug, Eg, 1.0, 3)
    Ut, Et = _gauge_transform(oracle, U, E, G)
    assert np.max(np.abs(Ug - Ut)) < 1e-12
    assert np.max(np.abs(Eg - Et)) < 1e-12
    assert oracle.energy(Ug, Eg) == pytest.approx(oracle.energy(U, E), rel=1e-13)
    # (the Gauss measure of GetGdeviation is a max over generator components
    # of g G g^+, which depends on the colour frame, so it is not compared)


def test_energy_error_is_second_order(oracle):
    """After a step the stored pair is U(t), E(t + dt/2) (links are updated
    first), so the energy of the stored pair fluctuates at O(dt).  With E
    synchronised to U, E(t) ~ (E(t - dt/2) + E(t + dt/2))/2, the leapfrog
    energy error is O(dt^2): halving StepWidth over the same physical time
    quarters the maximal deviation."""
    oracle.lattice_init((4, 4, 4), (0.2, 1.0, 1.0, 1.0))
    U0, E0 = oracle.box_init(1, 0.5, 1.0, 1.0)
    devs = []
    for sw, n in ((0.5, 20), (0.25, 40)):
        U, E = U0.copy(order="F"), E0.copy(order="F")
        oracle.update_links(U, E, sw)  # first half of step 1: U(dt), E(dt/2) = E0
        e_prev = E.copy(order="F")
        oracle.update_efield(U, E, sw)
        e_ref = oracle.energy(U, np.asfortranarray((e_prev + E) / 2))
        d = 0.0
        for _ in range(n):
            oracle.update_links(U, E, sw)
            e_prev = E.copy(order="F")
            oracle.update_efield(U, E, sw)
            d = max(d, abs(oracle.energy(U, np.asfortranarray((e_prev + E) / 2)) - e_ref))
        devs.append(d / e_ref)
    assert devs[0] < 0.01
    assert 3.0 < devs[0] / devs[1] < 5.0


def test_gauss_constraint_conserved(oracle):
    oracle.lattice_init((6, 6, 6), SP)
    U, E = oracle.box_init(2, 0.5, 1.0, 1.0)
    g0 = oracle.gauss_max(U, E)
    oracle.leapfrog(U, E, 1.0, 50)
    assert abs(oracle.gauss_max(U, E) - g0) < 1e-10 * max(1.0, g0)
    assert oracle.gauss_ref(U, E) < 1e-9 * oracle.energy(U, E)


# --------------------------------------------------------------------------
# NRQCD Crank-Nicolson
# --------------------------------------------------------------------------
def _dense(oracle, U, mass, c1, sw, herm):
    n = 6 * oracle.vol
    M = np.zeros((n, n), complex)
    for k in range(n):
        e = np.zeros(n, complex)
        e[k] = 1
        M[:, k] = oracle.apply_system(U, e, mass, c1, sw, herm)
    return M


def test_system_matrix_structure_on_tiny_lattice(oracle):
    """2^3 lattice (x+i == x-i, both hops add): S^+ is the Hermitian
    conjugate of S, K = (S - 1)/alpha is Hermitian and positive
    semi-definite with spectrum in [0, 6/a^2 (per direction 2/a^2)]."""
    oracle.lattice_init((2, 2, 2), (0.05, 1.0, 0.8, 1.2))
    U, E = oracle.box_init(1, 0.5, 1.0, 1.0)
    mass, c1, sw = 2.0, 1.0 + 0.0j, 1.0
    S = _dense(oracle, U, mass, c1, sw, False)
    Sd = _dense(oracle, U, mass, c1, sw, True)
    assert np.allclose(Sd, S.conj().T, atol=1e-15)
    alpha = -1j * (sw * 0.05 / 2) * c1 / mass
    K = (S - np.eye(S.shape[0])) / alpha
    assert np.allclose(K, K.conj().T, atol=1e-13)
    w = np.linalg.eigvalsh(K)
    assert w.min() > -1e-12
    assert w.max() <= 2 * (1 / 1.0 ** 2 + 1 / 0.8 ** 2 + 1 / 1.2 ** 2) + 1e-12


def test_cn_free_field_closed_form(oracle):
    """U = 1: each plane wave is multiplied per step by
    (1 - i k dt c1/2M)/(1 + i k dt c1/2M), k = sum_i 2 sin^2(p_i a_i/2)/a_i^2."""
    ext, sp = (6, 4, 5), (0.05, 1.0, 0.9, 1.1)
    oracle.lattice_init(ext, sp)
    U, _ = oracle.cold()
    Q = oracle.point_prop(1)
    A = oracle.point_prop(1)
    mass, nsteps = 2.0, 4
    for _ in range(nsteps):
        rc, st = oracle.cn_update(U, Q, A, mass, [1, 0, 0, 0], 1.0, 1e-13)
        assert rc == 0
    shape = ext[::-1]
    kx = [2 * np.sin(np.pi * np.arange(n) / n) ** 2 / a ** 2 for n, a in zip(ext, sp[1:])]
    K = kx[2][:, None, None] + kx[1][None, :, None] + kx[0][None, None, :]
    dt = sp[0]
    fq = ((1 - 0.5j * dt * K / mass) / (1 + 0.5j * dt * K / mass)) ** nsteps
    fa = ((1 + 0.5j * dt * K / mass) / (1 - 0.5j * dt * K / mass)) ** nsteps
    delta = np.zeros(shape)
    delta.flat[0] = 1 / np.sqrt(6)
    q = np.fft.ifftn(fq * np.fft.fftn(delta)).reshape(-1)
    a = np.fft.ifftn(fa * np.fft.fftn(delta)).reshape(-1)
    for d in range(6):
        assert np.allclose(Q[d, d, :], q, atol=1e-12)
        assert np.allclose(A[d, d, :], a, atol=1e-12)
    off = Q.copy()
    for d in range(6):
        off[d, d, :] = 0
    assert np.max(np.abs(off)) < 1e-14
    # free correlators: all four equal (i/V) sum_p fq^2
    corr = oracle.correlators(Q, A)
    want = 1j * np.mean(((1 - 0.5j * dt * K / mass) / (1 + 0.5j * dt * K / mass)) ** (2 * nsteps))
    for key, v in corr.items():
        assert abs(v - want) < 1e-11, key
    # the _gconjg variants see A^T = A here
    for key, v in oracle.correlators(Q, A, gconjg=True).items():
        assert abs(v - want) < 1e-11, key


def test_cn_norm_conservation_and_reversibility(oracle):
    oracle.lattice_init((4, 4, 4), SP)
    U, E = oracle.box_init(4, 0.5, 1.0, 1.0)
    Q0 = oracle.point_prop(5)
    A0 = oracle.point_prop(9)
    Q, A = Q0.copy(order="F"), A0.copy(order="F")
    for _ in range(4):
        rc, st = oracle.cn_update(U, Q, A, 2.0, [1, 0, 0, 0], 1.0, 1e-13)
        assert rc == 0 and st.solves == 12 and st.max_true_relres <= 1e-13
    assert oracle.prop_norm(Q) == pytest.approx(1.0, abs=1e-11)
    assert oracle.prop_norm(A) == pytest.approx(1.0, abs=1e-11)
    assert np.max(np.abs(Q - Q0)) > 1e-3
    for _ in range(4):
        oracle.cn_update(U, Q, A, 2.0, [1, 0, 0, 0], -1.0, 1e-13)
    assert np.max(np.abs(Q - Q0)) < 1e-11
    assert np.max(np.abs(A - A0)) < 1e-11


def test_cn_gauge_covariance(oracle):
    oracle.lattice_init((4, 3, 4), SP)
    rng = np.random.default_rng(5)
    U, E = oracle.box_init(5, 0.5, 1.0, 1.0)
    G = _random_su3(rng, oracle.vol)
    Ug, _ = _gauge_transform(oracle, U, E, G)
    Q = np.asfortranarray(rng.normal(size=(6, 6, oracle.vol)) + 1j * rng.normal(size=(6, 6, oracle.vol)))
    A = np.asfortranarray(rng.normal(size=(6, 6, oracle.vol)) + 1j * rng.normal(size=(6, 6, oracle.vol)))

    def lift(Qp):  # psi(x) -> (g(x) (x) 1_2) psi(x) in the C2CS ordering
        out = np.empty_like(Qp)
        for x in range(oracle.vol):
            out[:, :, x] = np.kron(G[x], np.eye(2)) @ Qp[:, :, x]
        return np.asfortranarray(out)

    Qg, Ag = lift(Q), lift(A)
    oracle.cn_update(U, Q, A, 1.5, [1, 0, 0, 0], 1.0, 1e-13)
    oracle.cn_update(Ug, Qg, Ag, 1.5, [1, 0, 0, 0], 1.0, 1e-13)
    assert np.max(np.abs(Qg - lift(Q))) < 1e-11 * np.max(np.abs(Q))
    assert np.max(np.abs(Ag - lift(A))) < 1e-11 * np.max(np.abs(A))


def test_cn_matches_dense_solve(oracle):
    """GMRES result equals the dense solution of S^+ psi' = S psi."""
    oracle.lattice_init((3, 2, 2), SP)
    U, E = oracle.box_init(6, 0.9, 1.0, 1.0)
    rng = np.random.default_rng(6)
    Q = np.asfortranarray(rng.normal(size=(6, 6, oracle.vol)) + 1j * rng.normal(size=(6, 6, oracle.vol)))
    A = np.asfortranarray(rng.normal(size=(6, 6, oracle.vol)) + 1j * rng.normal(size=(6, 6, oracle.vol)))
    mass, c1, sw = 0.7, 0.9 + 0.2j, 2.0
    S = _dense(oracle, U, mass, c1, sw, False)
    Sa = _dense(oracle, U, -mass, np.conj(c1), sw, False)
    want_q = np.stack([np.linalg.solve(S.conj().T, S @ Q[:, j, :].T.reshape(-1)) for j in range(6)], 1)
    want_a = np.stack([np.linalg.solve(Sa.conj().T, Sa @ A[:, j, :].T.reshape(-1)) for j in range(6)], 1)
    oracle.cn_update(U, Q, A, mass, [c1, 0, 0, 0], sw, 1e-14)
    got_q = np.stack([Q[:, j, :].T.reshape(-1) for j in range(6)], 1)
    got_a = np.stack([A[:, j, :].T.reshape(-1) for j in range(6)], 1)
    assert np.allclose(got_q, want_q, atol=1e-12)
    assert np.allclose(got_a, want_a, atol=1e-12)


def test_cn_rejects_higher_wilson_terms(oracle):
    oracle.lattice_init((2, 2, 2), SP)
    U, _ = oracle.cold()
    Q = oracle.point_prop(1)
    rc, _ = oracle.cn_update(U, Q, Q.copy(order="F"), 2.0, [1, 0.1, 0, 0], 1.0, 1e-10)
    assert rc == -3
