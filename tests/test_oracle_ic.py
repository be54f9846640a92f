"""Pinning the CPU restatement of the transverse-polarised initial conditions
with a general occupation (oracle/ic_oracle.py); no GPU."""
import numpy as np
import pytest

from oracle import gaugefix_oracle as gf
from oracle import ic_oracle as ico


def _setup(oracle, ext, sp, seed):
    oracle.lattice_init(ext, sp)
    oracle.ranlxd_init(2, seed)
    vol = int(np.prod(ext))
    pn = np.array([oracle.norm2_momentum(k + 1) for k in range(vol)])
    return vol, pn


def test_per_site_form_reproduces_box_ic(oracle):
    """With the box occupation the per-site form is the box initial condition
    of the C oracle."""
    ext, sp, seed, qs, amp, g = (4, 5, 3), (0.05, 1.0, 0.9, 1.1), 6, 0.7, 1.3, 0.8
    vol, pn = _setup(oracle, ext, sp, seed)
    occ = ico.box_occupation(pn, qs, amp, g)
    U, E, aa, ee, _, _ = ico.transverse_init(oracle, ext, sp, occ, False)
    U0, E0 = oracle.box_init(seed, qs, amp, g)
    assert np.max(np.abs(U - U0)) < 1e-13
    assert np.max(np.abs(E - E0)) < 1e-13
    assert aa[0] == 0 and ee[0] == 0 and np.all(aa >= -1e-15) and np.all(ee >= -1e-15)
    # The per-site form's momentum-space fields are not Hermitian (the second
    # polarisation vector obeys eps2(-p) = conj(eps2(p)), the first
    # eps1(-p) = -conj(eps1(p)), against the same factor i), so the
    # configuration keeps only part of them and the measured correlators of the
    # finished links differ from these, unlike the specific-sender form.
    lat = gf.Lattice(ext, sp, oracle.generators())
    _, mee = gf.transverse_correlators(lat, U, E)
    assert np.max(np.abs(ee - mee)) > 1e-2 * ee.max()


@pytest.mark.parametrize("ext", [(4, 4, 4), (5, 4, 3)])
def test_specific_sender_fields_are_real_and_transverse(oracle, ext):
    """The momentum-space symmetrisation makes the position-space fields real
    before their real part is taken, and the transverse polarisations give
    the lattice Coulomb condition sum_i (F_i(x) - F_i(x - i)) / a_i = 0."""
    sp = (0.05, 1.0, 1.0, 1.0)
    vol, pn = _setup(oracle, ext, sp, 3)
    occ = ico.box_occupation(pn, 0.8, 1.0, 1.0)
    U, E, aa, ee, Ax, Ex = ico.transverse_init(oracle, ext, sp, occ, True)
    assert np.max(np.abs(Ax.imag)) < 1e-13 * np.max(np.abs(Ax.real))
    assert np.max(np.abs(Ex.imag)) < 1e-13 * np.max(np.abs(Ex.real))
    div = np.zeros((vol, 8))
    for x in range(vol):
        for i in range(3):
            xm = oracle.neib(-(i + 1), x + 1) - 1
            div[x] += (Ex[x, i].real - Ex[xm, i].real) / sp[i + 1]
    assert np.max(np.abs(div)) < 1e-12 * np.max(np.abs(Ex.real))
    # the fields are real, so the correlators the measurement takes from the
    # finished configuration (group log of the links, E) are these
    lat = gf.Lattice(ext, sp, oracle.generators())
    maa, mee = gf.transverse_correlators(lat, U, E)
    assert np.allclose(aa, maa, rtol=1e-10, atol=1e-12 * aa.max())
    assert np.allclose(ee, mee, rtol=1e-10, atol=1e-12 * ee.max())
    # the momentum-space correlators are symmetric under p -> -p
    for k in range(vol):
        t = oracle.negative_index(k + 1) - 1
        assert aa[k] == pytest.approx(aa[t], rel=1e-12, abs=1e-300)


def test_specific_sender_single_mode_plane_wave(oracle):
    """Occupation f at one momentum index k0 (and at -k0), zero elsewhere.
    Lattice indices 2 .. k0-1 each draw 8 x 16 uniforms first (index 1 is
    p = 0 and draws none), so the amplitudes of generator a are the first two
    Box-Muller normals of the 16 uniforms at offset 128 (k0 - 2) + 16 a.  The
    field is the plane wave A_i(x) = 2 Re[A_i(k0) exp(i p_k0 . x)] / Vphys
    with A(k0) = sqrt(Vphys f / |p| / 2) (r1 eps1 + r2 eps2)."""
    ext, sp, seed = (6, 4, 3), (0.05, 1.0, 1.0, 1.0), 11
    vol, pn = _setup(oracle, ext, sp, seed)
    k0 = 1 + 2 + 6 * 1  # 0-based position (2, 1, 0), 1-based index 9; -k0 = (4, 3, 0)
    t0 = oracle.negative_index(k0)
    assert t0 > k0
    occ = np.zeros(vol)
    occ[k0 - 1] = occ[t0 - 1] = 0.7
    U, E, aa, ee, Ax, Ex = ico.transverse_init(oracle, ext, sp, occ, True)
    oracle.ranlxd_init(2, seed)
    u = oracle.ranlxd(128 * (k0 - 1))[128 * (k0 - 2):].reshape(8, 16)
    r1 = np.sqrt(-2 * np.log(u[:, 0])) * np.exp(2j * np.pi * u[:, 1])
    r2 = np.sqrt(-2 * np.log(u[:, 2])) * np.exp(2j * np.pi * u[:, 3])
    p = oracle.momentum(k0)
    eps = oracle.polarisation(p)
    vol_phys = vol * np.prod(sp[1:])
    Ak = np.sqrt(vol_phys * 0.7 / pn[k0 - 1] / 2) * (eps[:, 0][:, None] * r1[None, :] + eps[:, 1][:, None] * r2[None, :])
    n0 = np.array([2, 1, 0])
    for x in range(vol):
        pos = np.array([x % 6, (x // 6) % 4, x // 24])
        phase = np.exp(2j * np.pi * np.sum(n0 * pos / np.array(ext)))
        want = 2 * np.real(Ak * phase) / vol_phys  # (i, a)
        assert np.allclose(Ax[x].real, want, atol=1e-14)
        for i in range(3):
            assert np.allclose(U[:, :, i, x], oracle.group_exp(want[i] * sp[i + 1]), atol=1e-13)
    # only the occupied pair carries a correlator
    nz = np.flatnonzero(np.abs(aa) > 1e-14 * np.abs(aa).max())
    assert sorted(nz) == sorted([k0 - 1, t0 - 1])
