"""Pinning the CPU restatement of the full NRQCD system matrix
(oracle/hamiltonian_oracle.py, src/nrqcd.f90:1175-1387); no GPU."""
import numpy as np
import pytest

from oracle import gaugefix_oracle as gf
from oracle import hamiltonian_oracle as ho
from oracle.pyoracle import RefModules

SP = (0.05, 1.0, 1.0, 1.0)


def _lat(oracle, ext, sp=SP):
    oracle.lattice_init(ext, sp)
    return gf.Lattice(ext, sp, oracle.generators())


def _pauli(oracle):
    return [oracle.pauli(k) for k in (1, 2, 3)]


def _evolved(oracle, seed=4, steps=3):
    U, E = oracle.box_init(seed, 0.7, 1.0, 1.0)
    oracle.leapfrog(U, E, 1.0, steps)
    return U, E


def test_kinetic_part_equals_c_oracle(oracle):
    ext, sp = (4, 3, 5), (0.05, 1.0, 0.9, 1.2)
    lat = _lat(oracle, ext, sp)
    U, E = _evolved(oracle)
    c1 = 0.8 + 0.3j
    S = ho.system_matrix(lat, U, E, 1.7, [c1, 0, 0, 0], 0.6, _pauli(oracle))
    rng = np.random.default_rng(1)
    v = rng.normal(size=6 * lat.vol) + 1j * rng.normal(size=6 * lat.vol)
    for herm in (False, True):
        want = oracle.apply_system(U, v, 1.7, c1, 0.6, herm)
        got = (S.conj().T if herm else S) @ v
        assert np.max(np.abs(got - want)) < 1e-13 * np.max(np.abs(want))


@pytest.mark.skipif(not RefModules.available(), reason="oracle/_ref not built")
def test_arglogu_matches_compiled_reference(oracle):
    """Full matrix including the trace part, on plaquettes and on products
    whose principal logarithm is not traceless."""
    ref = RefModules()
    lat = _lat(oracle, (3, 3, 3))
    rng = np.random.default_rng(2)
    mats = [oracle.group_exp(rng.normal(size=8) * s) for s in (0.1, 1.0, 2.5) for _ in range(5)]
    mats += [oracle.group_exp(rng.normal(size=8) * 2.0) @ oracle.group_exp(rng.normal(size=8) * 2.0)
             for _ in range(5)]
    for m in mats:
        assert np.allclose(ho.arglogu(m[None])[0], ref.arglogu(m), atol=1e-11)


def test_c2_constant_abelian_magnetic_field(oracle):
    """U_x(x) = diag(exp(i theta_c y)), theta_c = 2 pi m_c / N_y, other links
    1: every xy plaquette is diag(exp(-i theta_c)), so F_xy = -F_yx =
    diag(-theta)/(a_x a_y), B_z = diag(-theta)/(a_x a_y), B_x = B_y = 0, and
    S - 1 (c = (0, c2, 0, 0)) is the site-diagonal block
    (-i dt a0/2)(-c2)/(2M) C2CS(B_z)."""
    ext, sp = (3, 5, 2), (0.05, 1.0, 0.9, 1.1)
    lat = _lat(oracle, ext, sp)
    U, E = oracle.cold()
    m = np.array([1, 1, -2])
    theta = 2 * np.pi * m / ext[1]
    for x in range(lat.vol):
        y = (x // ext[0]) % ext[1]
        U[:, :, 0, x] = np.diag(np.exp(1j * theta * y))
    c2, mass, dt = 0.7 - 0.2j, 1.5, 0.8
    S = ho.system_matrix(lat, U, E, mass, [0, c2, 0, 0], dt, _pauli(oracle))
    Bz = np.diag(-theta) / (sp[1] * sp[2])
    block = (-1j * dt / 2 * sp[0]) * (-c2) / 2 / mass * np.kron(Bz, np.eye(2))
    want = np.kron(np.eye(lat.vol), np.eye(6) + block)
    assert np.max(np.abs(S.toarray() - want)) < 1e-13


def test_c3_against_algebra_matrices(oracle):
    """U = 1: D_i(x) = (e_i(x+i) - e_i(x-i))/(2 a_i) with
    e_i = -Re(E_i^a T^a)/(a0 a_i), the real part taken element by element; the
    exponential/logarithm round trip is the identity for small E."""
    ext, sp = (4, 3, 3), (0.05, 1.0, 0.8, 1.3)
    lat = _lat(oracle, ext, sp)
    U, E = oracle.cold()
    rng = np.random.default_rng(3)
    E[:] = rng.normal(size=E.shape) * 0.3
    c3, mass, dt = 0.9 + 0.4j, 2.0, 1.0
    S = ho.system_matrix(lat, U, E, mass, [0, 0, c3, 0], dt, _pauli(oracle))
    T = lat.T
    e = [-np.real(np.einsum("xa,rca->xrc", E[:, i, :].T, T)) / (sp[0] * sp[i + 1]) for i in range(3)]
    pre = -1j * dt / 2 * sp[0] * (-c3) / 8 / mass ** 2
    for x in range(lat.vol):
        D = sum((e[i][lat.up[i][x]] - e[i][lat.dn[i][x]]) / (2 * sp[i + 1]) for i in range(3))
        blk = S[6 * x:6 * x + 6, 6 * x:6 * x + 6].toarray()
        assert np.max(np.abs(blk - np.eye(6) - pre * np.kron(D, np.eye(2)))) < 1e-13


def test_c4_plane_wave_multiplier(oracle):
    """U = 1 and the same E at every site: the c4 hops are translation
    invariant, and on psi(x) = v exp(i p.x) the operator S - 1 multiplies by
    sum_i sum_(j,k,eps) (-i dt a0/2) eps c4/(8 M^2) i (exp(-i p_i) - exp(i p_i))
    Kron(e_j / a_i, sigma_k)  (lattice momenta p_i = 2 pi n_i / N_i)."""
    ext, sp = (4, 5, 3), (0.05, 1.0, 0.9, 1.1)
    lat = _lat(oracle, ext, sp)
    U, E = oracle.cold()
    rng = np.random.default_rng(4)
    e0 = rng.normal(size=(8, 3)) * 0.4
    for x in range(lat.vol):
        E[:, :, x] = e0
    c4, mass, dt = 0.6 - 0.5j, 1.3, 0.9
    pauli = _pauli(oracle)
    S = ho.system_matrix(lat, U, E, mass, [0, 0, 0, c4], dt, pauli)
    T = lat.T
    e = [-np.real(np.einsum("a,rca->rc", e0[:, j], T)) / (sp[0] * sp[j + 1]) for j in range(3)]
    n = np.array([1, 2, 2])
    pos = np.stack([np.arange(lat.vol) % ext[0], (np.arange(lat.vol) // ext[0]) % ext[1],
                    np.arange(lat.vol) // (ext[0] * ext[1])], 1)
    phase = np.exp(2j * np.pi * (pos @ (n / np.array(ext))))
    v = rng.normal(size=6) + 1j * rng.normal(size=6)
    psi = (phase[:, None] * v[None, :]).reshape(-1)
    mult = np.zeros((6, 6), complex)
    for i in range(3):
        pi_ = 2 * np.pi * n[i] / ext[i]
        for (j, k, eps) in (((i + 1) % 3, (i + 2) % 3, 1), ((i + 2) % 3, (i + 1) % 3, -1)):
            mult += (-1j * dt / 2 * sp[0]) * eps * c4 / 8 / mass ** 2 * 1j * (np.exp(-1j * pi_) - np.exp(1j * pi_)) \
                * np.kron(e[j] / sp[i + 1], pauli[k])
    got = (S @ psi - psi).reshape(-1, 6)
    want = phase[:, None] * (mult @ v)[None, :]
    assert np.max(np.abs(got - want)) < 1e-13 * np.max(np.abs(want))


def test_cn_step_full_hamiltonian_keeps_the_source_structure(oracle):
    """With c4 = 0 the operator is colour-only, so spin rows of a column do not
    mix; with c4 != 0 they do.  The solves are exact: S^+ psi' = S psi."""
    ext = (3, 3, 4)
    lat = _lat(oracle, ext)
    U, E = _evolved(oracle, 6, 2)
    pauli = _pauli(oracle)
    Q, A = oracle.point_prop(1), oracle.point_prop(1)
    ho.cn_update(lat, U, E, Q, A, 2.0, [1, 1, 1, 0], 1.0, pauli)
    spin_mix = lambda P: max(np.abs(P[d, j, :]).max() for j in range(6) for d in range(6) if d % 2 != j % 2)
    assert spin_mix(Q) < 1e-14 and spin_mix(A) < 1e-14
    Q, A = oracle.point_prop(1), oracle.point_prop(1)
    ho.cn_update(lat, U, E, Q, A, 2.0, [1, 1, 1, 1], 1.0, pauli)
    assert spin_mix(Q) > 1e-8 and spin_mix(A) > 1e-8
    S = ho.system_matrix(lat, U, E, 2.0, [1, 1, 1, 1], 1.0, pauli)
    q0 = ho.prop_to_vec(oracle.point_prop(1), 0)
    assert np.max(np.abs(S.conj().T @ ho.prop_to_vec(Q, 0) - S @ q0)) < 1e-13
