"""The oracle's algebra, FFT and initial condition against mathematics that
fixes them independently of the oracle's own code: scipy/numpy routines,
closed forms and the exact transversality the construction guarantees."""
import numpy as np
import pytest
import scipy.linalg


def test_group_exp_is_expm(oracle):
    T = oracle.generators()
    rng = np.random.default_rng(1)
    for scale in (1e-8, 1e-3, 0.1, 1.0, 5.0):
        for _ in range(10):
            a = rng.normal(size=8) * scale
            H = np.einsum("a,rca->rc", a, T)
            want = scipy.linalg.expm(1j * H)
            got = oracle.group_exp(a)
            assert np.max(np.abs(got - want)) < 1e-14 * max(1.0, np.abs(a).sum())
            assert np.max(np.abs(got @ got.conj().T - np.eye(3))) < 1e-14
            assert abs(np.linalg.det(got) - 1) < 1e-13


def test_generators_are_su3_basis(oracle):
    T = oracle.generators()
    for a in range(8):
        assert np.allclose(T[:, :, a], T[:, :, a].conj().T)
        assert abs(np.trace(T[:, :, a])) < 1e-15
        for b in range(8):
            assert np.trace(T[:, :, a] @ T[:, :, b]) == pytest.approx(0.5 if a == b else 0.0, abs=1e-15)
    rng = np.random.default_rng(2)
    for _ in range(10):
        a = rng.normal(size=8)
        M = oracle.algebra_matrix(a)
        got = [oracle.algebra_coordinate(k, M) for k in range(1, 9)]
        assert np.allclose(got, a, atol=1e-14)


def test_fft_conventions(oracle):
    ext, sp = (4, 6, 5), (0.05, 1.0, 0.8, 1.2)
    oracle.lattice_init(ext, sp)
    rng = np.random.default_rng(3)
    x = rng.normal(size=oracle.vol) + 1j * rng.normal(size=oracle.vol)
    cube = x.reshape(ext[::-1])  # (z, y, x) with x fastest
    want = np.fft.fftn(cube).reshape(-1) * np.prod(sp[1:])
    got = oracle.x2p(x)
    assert np.allclose(got, want, rtol=1e-13, atol=1e-13)
    vol_phys = np.prod(ext) * np.prod(sp[1:])
    want_b = np.fft.ifftn(cube).reshape(-1) * np.prod(ext) / vol_phys
    assert np.allclose(oracle.p2x(x), want_b, rtol=1e-13, atol=1e-13)
    assert np.allclose(oracle.p2x(oracle.x2p(x)), x, atol=1e-13)


def test_polarisation_vectors_transverse_and_normalised(oracle):
    ext, sp = (6, 5, 4), (0.1, 1.0, 0.9, 1.1)
    oracle.lattice_init(ext, sp)
    for idx in range(2, oracle.vol + 1):
        p = oracle.momentum(idx)
        eps = oracle.polarisation(p)
        for pol in (0, 1):
            e = eps[:, pol]
            assert abs(np.sum(e * p)) < 1e-14  # bilinear transversality p . eps = 0
            assert np.linalg.norm(e) == pytest.approx(1.0, abs=1e-14)
        assert abs(np.vdot(eps[:, 0], eps[:, 1])) < 1e-14
        assert np.allclose(eps[:, 2], p / np.linalg.norm(np.abs(p)), atol=1e-15)
    assert np.all(oracle.polarisation(oracle.momentum(1)) == 0)


def test_polarisation_vectors_cross_product_form(oracle):
    """The polarisation vectors of lattice.f90:1120-1176 (after Kaspar and
    Hebenstreit, arXiv:1403.4849) by a second route, vector algebra: with
    p^ = p / |p|, eps1 = z^ x p / |(p1, p2)| when |p1| > 0, else
    eps1 = -x^ x p / |(p2, p3)|; eps2 = conj(eps1) x p^ and eps3 = p^.  This
    fixes what transversality and norms leave open (signs, where the complex
    conjugate sits, the order of the two transverse vectors, the branch)."""
    rng = np.random.default_rng(5)
    xh, zh = np.array([1, 0, 0], complex), np.array([0, 0, 1], complex)
    oracle.lattice_init((6, 5, 4), (0.1, 1.0, 0.9, 1.1))
    cases = [oracle.momentum(i) for i in range(2, oracle.vol + 1)]  # lattice momenta, both branches
    cases += [rng.normal(size=3) + 1j * rng.normal(size=3) for _ in range(50)]
    cases += [np.array([0, 0.3 - 0.2j, 1.1 + 0.4j]), np.array([0, 0, 0.7j]), np.array([0, -0.5, 0]), np.array([2.0, 0, 0])]
    branches = set()
    for p in cases:
        if np.linalg.norm(np.abs(p)) < 1e-12:
            continue
        ph = p / np.linalg.norm(np.abs(p))
        if abs(p[0]) > 1e-12:
            e1 = np.cross(zh, p) / np.linalg.norm(np.abs(p[:2]))
            branches.add(1)
        else:
            e1 = -np.cross(xh, p) / np.linalg.norm(np.abs(p[1:]))
            branches.add(2)
        e2 = np.cross(np.conj(e1), ph)
        eps = oracle.polarisation(p)
        assert np.allclose(eps[:, 0], e1, rtol=0, atol=1e-14), p
        assert np.allclose(eps[:, 1], e2, rtol=0, atol=1e-14), p
        assert np.allclose(eps[:, 2], ph, rtol=0, atol=1e-15), p
    assert branches == {1, 2}


def _algebra_from_link(U, T):
    """alpha with U = exp(i alpha^a T^a) (principal branch)."""
    w, V = np.linalg.eig(U)
    H = (V * np.angle(w)) @ np.linalg.inv(V)
    return np.array([2 * np.real(np.trace(T[:, :, a] @ H)) for a in range(8)])


@pytest.mark.parametrize("ext,sp", [((6, 6, 6), (0.05, 1.0, 1.0, 1.0)), ((4, 6, 5), (0.1, 1.0, 0.8, 1.2))])
def test_box_ic_is_transverse(oracle, ext, sp):
    """A(p) and E(p) are combinations of transverse polarisation vectors
    (gaugeconfiguration_su3.f90:1353-1381), so the fields satisfy the lattice
    Coulomb condition sum_i (F_i(x) - F_i(x - i))/a_i = 0 with F = E/(a0 a_i)
    and F = Re A = alpha/a_i exactly (up to rounding)."""
    oracle.lattice_init(ext, sp)
    U, E = oracle.box_init(5, 0.8, 1.0, 1.0)
    T = oracle.generators()
    V = oracle.vol
    a0, a = sp[0], sp[1:]
    alpha = np.zeros((8, 3, V))
    for x in range(V):
        for i in range(3):
            alpha[:, i, x] = _algebra_from_link(U[:, :, i, x], T)
    divE = np.zeros((8, V))
    divA = np.zeros((8, V))
    for x in range(V):
        for i in range(3):
            xm = oracle.neib(-(i + 1), x + 1) - 1
            divE[:, x] += (E[:, i, x] - E[:, i, xm]) / (a0 * a[i]) / a[i]
            divA[:, x] += (alpha[:, i, x] - alpha[:, i, xm]) / a[i] / a[i]
    assert np.max(np.abs(E)) > 1e-3
    assert np.max(np.abs(divE)) < 1e-12 * np.max(np.abs(E)) / a0
    assert np.max(np.abs(divA)) < 1e-11


def test_box_ic_restated_in_numpy(oracle):
    """Independent numpy assembly of the box IC from the pinned ranlxd stream
    (numpy FFT, numpy polarisation algebra)."""
    ext, sp, seed, qs, amp, g = (4, 4, 4), (0.05, 1.0, 1.0, 1.0), 3, 0.6, 1.2, 0.9
    oracle.lattice_init(ext, sp)
    U, E = oracle.box_init(seed, qs, amp, g)
    oracle.ranlxd_init(2, seed)
    V = oracle.vol
    u = oracle.ranlxd(32 * V).reshape(V, 2, 16)
    r = np.sqrt(-2 * np.log(u[:, :, 0::2])) * np.cos(2 * np.pi * u[:, :, 1::2])  # (V, pol, a)
    shape = ext[::-1]
    rk = np.fft.fftn(r.transpose(1, 2, 0).reshape(2, 8, *shape), axes=(2, 3, 4)).reshape(2, 8, V) * np.prod(sp[1:])
    A = np.zeros((3, 8, V), complex)
    Ef = np.zeros((3, 8, V), complex)
    for x in range(V):
        p = oracle.momentum(x + 1)
        pn = np.linalg.norm(np.abs(p))
        if pn <= 1e-12:
            continue
        ab = np.abs(p)
        if ab[0] > 1e-12:
            n12 = np.hypot(ab[0], ab[1])
            e1 = np.array([-p[1], p[0], 0]) / n12
            e2 = np.array([np.conj(p[0]) * p[2], np.conj(p[1]) * p[2], -n12 ** 2]) / (pn * n12)
        else:
            n23 = np.hypot(ab[1], ab[2])
            e1 = np.array([0, p[2], -p[1]]) / n23
            e2 = np.array([1, 0, 0])
        f = amp if pn < qs else g * g / 2
        mode = rk[0, :, x][None, :] * e1[:, None] + rk[1, :, x][None, :] * e2[:, None]
        A[:, :, x] = np.sqrt(2 * f / pn) * mode
        Ef[:, :, x] = 1j * np.sqrt(2 * f * pn) * mode
    vol_phys = np.prod(ext) * np.prod(sp[1:])
    Ax = np.fft.ifftn(A.reshape(3, 8, *shape), axes=(2, 3, 4)).reshape(3, 8, V) * V / vol_phys
    Ex = np.fft.ifftn(Ef.reshape(3, 8, *shape), axes=(2, 3, 4)).reshape(3, 8, V) * V / vol_phys
    T = oracle.generators()
    for x in range(V):
        for i in range(3):
            H = np.einsum("a,rca->rc", Ax[i, :, x].real * sp[i + 1], T)
            assert np.allclose(U[:, :, i, x], scipy.linalg.expm(1j * H), atol=1e-13)
            assert np.allclose(E[:, i, x], Ex[i, :, x].real * sp[0] * sp[i + 1], atol=1e-13)
