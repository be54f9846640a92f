"""The Fortran drop-in (nrqcd_amd/fortran: modules lattice, random,
gaugeconfiguration_su3, nrqcd and the nrqcd_run driver with the reference's
positional command line) run end to end on the GPU and checked against the
oracle."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "nrqcd_amd" / "fortran" / "nrqcd_run"


def _run(args, cwd):
    print("nrqcd_run", " ".join(map(str, args)), flush=True)
    if not EXE.exists():
        pytest.fail("nrqcd_run not built (python -m nrqcd_amd.build)")
    r = subprocess.run([str(EXE), *map(str, args)], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


def test_mode2_energy_and_gauss(oracle, tmp_path):
    ext, sp, seed = (6, 6, 6), (0.05, 1.0, 1.0, 1.0), 3
    # mode Nx Ny Nz a0 ax ay az CoMTime TimeRange Seed Qs Amp g
    _run([2, *ext, *sp, 0.0, 0.5, seed, 0.5, 1.0, 1.0], tmp_path)
    rows = np.loadtxt(tmp_path / "energy_gauss_constraint_plaquette.txt")
    short = np.loadtxt(tmp_path / "energy_gauss.txt")
    assert rows.shape == (11, 5) and short.shape == (11, 3)
    oracle.lattice_init(ext, sp)
    U, E = oracle.box_init(seed, 0.5, 1.0, 1.0)
    for k in range(rows.shape[0]):
        assert rows[k, 0] == pytest.approx(k * sp[0], abs=1e-15)
        assert rows[k, 1] == pytest.approx(oracle.energy(U, E), rel=1e-12)
        assert rows[k, 3] == pytest.approx(oracle.gauss_max(U, E), rel=1e-10)
        assert rows[k, 4] == pytest.approx(oracle.mean_plaquette(U), rel=1e-12)
        assert abs(rows[k, 2]) < 1e-8 * rows[k, 1]
        assert short[k, 1] == pytest.approx(rows[k, 1], rel=1e-5)
        oracle.leapfrog(U, E, 1.0, 1)


def test_mode3_correlator_grid(oracle, tmp_path):
    ext, seed, ds, smax = (6, 6, 6), 2, 0.1, 0.2
    a0 = ds / 2
    mass, tol = 2.0, 1e-12
    # mode Nx Ny Nz ds ax ay az tmin tmax smax Seed Qs Amp g kspTol M c1..c4 (re, im) files
    _run([3, *ext, ds, 1.0, 1.0, 1.0, 0.0, 0.0, smax, seed, 0.5, 1.0, 1.0, tol, mass,
          1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, "corr.txt", "norm.txt"], tmp_path)
    corr = np.loadtxt(tmp_path / "corr.txt", skiprows=1)
    norm = np.loadtxt(tmp_path / "norm.txt", skiprows=1)
    assert corr.shape == (3, 3)
    assert np.allclose(corr[:, 0], [0.0, ds, 2 * ds])
    assert corr[0, 1] == 0.0 and corr[0, 2] == 1.0
    # oracle replay of main.f90:2275-2364 for tmin = tmax = 0, smax = 2 ds:
    # t1 = -2 a0 ... 0; (s=2ds) from t1 = -2 a0 after 4 CN+leapfrog steps,
    # (s=ds) from t1 = -a0 after 2 steps
    sp = (a0, 1.0, 1.0, 1.0)
    oracle.lattice_init(ext, sp)
    U, E = oracle.box_init(seed, 0.5, 1.0, 1.0)
    oracle.leapfrog(U, E, -1.0, 2)
    want = {}
    for it1, nsteps, srow in ((-2, 4, 2), (-1, 2, 1)):
        U2, E2 = U.copy(order="F"), E.copy(order="F")
        Q, A = oracle.point_prop(1), oracle.point_prop(1)
        for _ in range(nsteps):
            rc, _ = oracle.cn_update(U2, Q, A, mass, [1, 0, 0, 0], 1.0, tol)
            assert rc == 0
            oracle.leapfrog(U2, E2, 1.0, 1)
        want[srow] = (oracle.correlators(Q, A)["O1_3S1"], oracle.prop_norm(Q), oracle.prop_norm(A))
        oracle.leapfrog(U, E, 1.0, 1)
    for srow, (c, nq, na) in want.items():
        assert corr[srow, 1] == pytest.approx(c.real, rel=1e-8, abs=1e-9)
        assert corr[srow, 2] == pytest.approx(c.imag, rel=1e-8, abs=1e-9)
        assert norm[srow, 1] == pytest.approx(nq, rel=1e-8)
        assert norm[srow, 2] == pytest.approx(na, rel=1e-8)


def test_mode1_gluon_distribution(oracle, tmp_path):
    """main.f90:3020-3190: box IC evolved to CoMTime, Coulomb gauge fixing of a
    copy, |p| and sqrt(AA EE) per momentum, against the oracle's replay."""
    from oracle import gaugefix_oracle as gf

    ext, sp, seed = (8, 6, 5), (0.05, 1.0, 1.0, 1.0), 4
    comtime, maxit, tol, alpha = 0.25, 500, 1e-10, 0.08
    # mode Nx Ny Nz a0 ax ay az CoMTime Seed Qs Amp g GFMaxIt GFTol GFCoeff File
    _run([1, *ext, *sp, comtime, seed, 0.5, 1.0, 1.0, maxit, tol, alpha, "gluon.txt"], tmp_path)
    rows = np.loadtxt(tmp_path / "gluon.txt")
    oracle.lattice_init(ext, sp)
    lat = gf.Lattice(ext, sp, oracle.generators())
    assert rows.shape == (lat.vol, 2)
    U, E = oracle.box_init(seed, 0.5, 1.0, 1.0)
    oracle.leapfrog(U, E, 1.0, int(np.ceil(comtime / sp[0])))
    V, _, dev = gf.coulomb_gauge_fix(lat, U, tol, alpha, maxit)
    assert dev <= tol
    aa, ee = gf.transverse_correlators(lat, V, E)
    f = np.sqrt(aa * ee)
    # E13.6 keeps 6 significant digits: relative rounding up to 5e-6
    assert np.allclose(rows[:, 0], lat.pnorm, rtol=5e-6, atol=1e-12)
    assert np.allclose(rows[:, 1], f, rtol=5e-6, atol=1e-6 * f.max())


KEYS = ("O1_1S0", "O1_3S1", "O8_1S0", "O8_3S1")
QUARK_ARGS = (1e-12, 2.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)  # kspTol, M, c1..c4 (re, im)


def _correlator_row(c):
    return np.array([v for k in KEYS for v in (c[k].real, c[k].imag)])


def test_mode32_free_correlators(oracle, tmp_path):
    """main.f90:282-522 in a cold gauge field: every row against the free
    closed form, all four channels equal i <f^(2n)>_p with the Crank-Nicolson
    phase f = (1 - i dt K c1/2M)/(1 + i dt K c1/2M), norms 1."""
    ext, sp, smax, mass = (6, 4, 5), (0.05, 1.0, 0.9, 1.1), 0.2, 2.0
    # mode Nx Ny Nz a0 ax ay az smax Seed kspTol M c1..c4 FileCorrelator FileNorm
    _run([32, *ext, *sp, smax, 1, *QUARK_ARGS, "corr.txt", "norm.txt"], tmp_path)
    header = (tmp_path / "corr.txt").read_text().splitlines()[0]
    assert header.split() == ["s", "Re(O1(1S0))", "Im(O1(1S0))", "Re(O1(3S1))", "Im(O1(3S1))",
                              "Re(O8(1S0))", "Im(O8(1S0))", "Re(O8(3S1))", "Im(O8(3S1))"]
    corr = np.loadtxt(tmp_path / "corr.txt", skiprows=1)
    norm = np.loadtxt(tmp_path / "norm.txt", skiprows=1)
    assert corr.shape == (5, 9) and norm.shape == (5, 3)
    kx = [2 * np.sin(np.pi * np.arange(n) / n) ** 2 / a ** 2 for n, a in zip(ext, sp[1:])]
    K = kx[2][:, None, None] + kx[1][None, :, None] + kx[0][None, None, :]
    f = (1 - 0.5j * sp[0] * K / mass) / (1 + 0.5j * sp[0] * K / mass)
    for n in range(5):
        want = 1j * np.mean(f ** (2 * n))
        assert corr[n, 0] == pytest.approx(n * sp[0], abs=1e-12)
        for ch in range(4):
            assert corr[n, 1 + 2 * ch] == pytest.approx(want.real, rel=1e-8, abs=1e-9)
            assert corr[n, 2 + 2 * ch] == pytest.approx(want.imag, rel=1e-8, abs=1e-9)
        assert norm[n, 1] == pytest.approx(1.0, rel=1e-9) and norm[n, 2] == pytest.approx(1.0, rel=1e-9)


def test_mode31_correlators_at_one_time(oracle, tmp_path):
    """main.f90:524-830 against an oracle replay: box IC evolved to t; for
    it1 = 1, 2 the point source and a copy of the gauge field are evolved
    2 it1 steps (CN, then the leapfrog); the gauge field then steps back once.
    All four correlators, their _gconjg variants and the norms."""
    ext, sp, seed = (6, 5, 4), (0.05, 1.0, 1.0, 1.0), 5
    t, smax, mass, tol = 0.1, 0.2, 2.0, 1e-12
    # mode Nx Ny Nz a0 ax ay az t smax Seed Qs Amp g kspTol M c1..c4 files
    _run([31, *ext, *sp, t, smax, seed, 0.5, 1.0, 1.0, *QUARK_ARGS, "corr.txt", "norm.txt", "corr_g.txt"],
         tmp_path)
    corr = np.loadtxt(tmp_path / "corr.txt", skiprows=1)
    corr_g = np.loadtxt(tmp_path / "corr_g.txt", skiprows=1)
    norm = np.loadtxt(tmp_path / "norm.txt", skiprows=1)
    assert corr.shape == (3, 9) and corr_g.shape == (3, 9) and norm.shape == (3, 3)
    oracle.lattice_init(ext, sp)
    U, E = oracle.box_init(seed, 0.5, 1.0, 1.0)
    oracle.leapfrog(U, E, 1.0, 2)
    rows, rows_g, norms = [], [], []
    Q, A = oracle.point_prop(1), oracle.point_prop(1)
    rows.append(_correlator_row(oracle.correlators(Q, A)))
    rows_g.append(_correlator_row(oracle.correlators(Q, A, gconjg=True)))
    norms.append((oracle.prop_norm(Q), oracle.prop_norm(A)))
    for it1 in (1, 2):
        U2, E2 = U.copy(order="F"), E.copy(order="F")
        Q, A = oracle.point_prop(1), oracle.point_prop(1)
        for _ in range(2 * it1):
            rc, _ = oracle.cn_update(U2, Q, A, mass, [1, 0, 0, 0], 1.0, tol)
            assert rc == 0
            oracle.leapfrog(U2, E2, 1.0, 1)
        rows.append(_correlator_row(oracle.correlators(Q, A)))
        rows_g.append(_correlator_row(oracle.correlators(Q, A, gconjg=True)))
        norms.append((oracle.prop_norm(Q), oracle.prop_norm(A)))
        oracle.leapfrog(U, E, -1.0, 1)
    assert np.allclose(corr[:, 0], [0.0, 2 * sp[0], 4 * sp[0]], atol=1e-12)
    # E16.9 keeps 10 significant digits
    assert np.allclose(corr[:, 1:], np.array(rows), rtol=1e-8, atol=1e-9)
    assert np.allclose(corr_g[:, 1:], np.array(rows_g), rtol=1e-8, atol=1e-9)
    assert np.allclose(norm[:, 1:], np.array(norms), rtol=1e-8)
    # the two correlator sets differ once the gauge field has acted
    assert np.max(np.abs(corr[2, 1:] - corr_g[2, 1:])) > 1e-6


def test_mode4_spectrum(tmp_path):
    """main.f90:1865-2100 on a four-channel correlator file: the first
    channel up to MaxTime, its spectrum with the frequencies in writing order
    and the intermediate signals, against the oracle's replay."""
    from oracle import spectrum_oracle as so

    dt, nrows, maxtime = 0.1, 60, 4.0
    t = np.arange(nrows) * dt
    c = np.exp(-1j * 1.3 * t) * np.exp(-0.05 * t) + 0.3 * np.exp(-1j * 4.1 * t)
    with open(tmp_path / "mesoncorrelator.txt", "w") as fh:
        fh.write("s Re(O1(1S0)) Im(O1(1S0)) Re(O1(3S1)) Im(O1(3S1))\n")
        for k in range(nrows):
            fh.write(f"{t[k]:+.9E} {c[k].real:+.9E} {c[k].imag:+.9E} {0.0:+.9E} {1.0:+.9E}\n")
    # mode nHeaderLines FileInput FileOutput MaxTime
    _run([4, 1, "mesoncorrelator.txt", "spectrum.txt", maxtime], tmp_path)
    sig, sdt = so.read_signal(tmp_path / "mesoncorrelator.txt", 1, maxtime)
    assert sig.size == 41
    spec, (s, h, f) = so.spectrum(sig, sdt)
    rows = np.loadtxt(tmp_path / "spectrum.txt")
    want = so.write_order(spec, sdt)
    assert rows.shape == want.shape == (80, 2)
    # E13.6 keeps 6 significant digits
    assert np.allclose(rows[:, 0], want[:, 0], rtol=5e-6, atol=0)
    assert np.allclose(rows[:, 1], want[:, 1], rtol=5e-6, atol=1e-6 * np.abs(want[:, 1]).max())
    for name, x in (("input.txt", sig), ("symm.txt", s), ("hanned.txt", h), ("fftformat.txt", f)):
        got = np.loadtxt(tmp_path / name)
        assert np.array_equal(got[:, 0], np.arange(1, x.size + 1))
        assert np.allclose(got[:, 1] + 1j * got[:, 2], x, rtol=1e-12, atol=1e-14), name


def test_mode31_with_all_wilson_coefficients(oracle, tmp_path):
    """The reference's templates run with c = (1, 1, 1, 1): mode 31 with the
    full Hamiltonian against a replay with oracle/hamiltonian_oracle.py."""
    from oracle import gaugefix_oracle as gf
    from oracle import hamiltonian_oracle as ho

    ext, sp, seed = (4, 3, 5), (0.05, 1.0, 1.0, 1.0), 8
    t, smax, mass = 0.05, 0.2, 2.0
    quark = (1e-12, mass, 1.0, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0, 0.0)
    _run([31, *ext, *sp, t, smax, seed, 0.5, 1.0, 1.0, *quark, "corr.txt", "norm.txt", "corr_g.txt"], tmp_path)
    corr = np.loadtxt(tmp_path / "corr.txt", skiprows=1)
    corr_g = np.loadtxt(tmp_path / "corr_g.txt", skiprows=1)
    norm = np.loadtxt(tmp_path / "norm.txt", skiprows=1)
    oracle.lattice_init(ext, sp)
    lat = gf.Lattice(ext, sp, oracle.generators())
    pauli = [oracle.pauli(k) for k in (1, 2, 3)]
    U, E = oracle.box_init(seed, 0.5, 1.0, 1.0)
    oracle.leapfrog(U, E, 1.0, 1)
    rows, rows_g, norms = [], [], []
    for it1 in (0, 1, 2):
        U2, E2 = U.copy(order="F"), E.copy(order="F")
        Q, A = oracle.point_prop(1), oracle.point_prop(1)
        for _ in range(2 * it1):
            ho.cn_update(lat, U2, E2, Q, A, mass, [1, 1, 1, 1], 1.0, pauli)
            oracle.leapfrog(U2, E2, 1.0, 1)
        rows.append(_correlator_row(oracle.correlators(Q, A)))
        rows_g.append(_correlator_row(oracle.correlators(Q, A, gconjg=True)))
        norms.append((oracle.prop_norm(Q), oracle.prop_norm(A)))
        if it1 > 0:
            oracle.leapfrog(U, E, -1.0, 1)
    assert np.allclose(corr[:, 1:], np.array(rows), rtol=1e-8, atol=1e-9)
    assert np.allclose(corr_g[:, 1:], np.array(rows_g), rtol=1e-8, atol=1e-9)
    assert np.allclose(norm[:, 1:], np.array(norms), rtol=1e-8)


def test_mode33_equilibrium_correlators(oracle, tmp_path):
    """main.f90:18-280: EquilibriumInit, then the point source with CN and
    leapfrog steps and all four correlators per step; the energy file of the
    optional MeasureEnergy flag.  Against an oracle replay."""
    from oracle import equilibrium_oracle as eo
    from oracle import gaugefix_oracle as gf

    ext, sp, seed = (4, 3, 4), (0.05, 1.0, 1.0, 1.0), 6
    beta, nef, neq, smax = 16.0, 1, 3, 0.1
    # mode Nx Ny Nz a0 ax ay az smax Seed Beta nefieldinit nequilibrium kspTol M c1..c4 files flag file
    _run([33, *ext, *sp, smax, seed, beta, nef, neq, *QUARK_ARGS, "corr.txt", "norm.txt", 1, "energy.txt"], tmp_path)
    corr = np.loadtxt(tmp_path / "corr.txt", skiprows=1)
    norm = np.loadtxt(tmp_path / "norm.txt", skiprows=1)
    energy = np.loadtxt(tmp_path / "energy.txt", ndmin=2)
    oracle.lattice_init(ext, sp)
    lat = gf.Lattice(ext, sp, oracle.generators())
    oracle.ranlxd_init(2, seed)
    U, E, rec = eo.equilibrium_init(oracle, lat, beta, nef, neq)
    assert energy.shape == (1, 2) and energy[0, 0] == 1
    assert energy[0, 1] == pytest.approx(rec[0][1], rel=1e-10)
    Q, A = oracle.point_prop(1), oracle.point_prop(1)
    rows = [_correlator_row(oracle.correlators(Q, A))]
    norms = [(oracle.prop_norm(Q), oracle.prop_norm(A))]
    for _ in range(2):
        rc, _ = oracle.cn_update(U, Q, A, 2.0, [1, 0, 0, 0], 1.0, 1e-12)
        assert rc == 0
        oracle.leapfrog(U, E, 1.0, 1)
        rows.append(_correlator_row(oracle.correlators(Q, A)))
        norms.append((oracle.prop_norm(Q), oracle.prop_norm(A)))
    assert corr.shape == (3, 9)
    assert np.allclose(corr[:, 1:], np.array(rows), rtol=1e-8, atol=1e-9)
    assert np.allclose(norm[:, 1:], np.array(norms), rtol=1e-8)
