"""Ensemble statistics of the GPU evolution against the oracle's after the
dynamics have decorrelated.

Classical Yang-Mills evolution is chaotic: after 200 steps (t = 10 a, twice
the 1/Qs time scale) two FP64 implementations that start from the same
seeded box initial condition no longer agree link by link to the 1e-12 of
the 10-step parity tests.  What must agree are ensemble averages
(gaugeconfiguration_su3.f90:1722-1737, 1769-1817; main.f90:3093-3189):
the energy density, the mean plaquette and the gluon occupation
f(p) = sqrt(<AA(p)> <EE(p)>) after Coulomb gauge fixing, averaged in |p| bins.

Two comparisons at 16^3 (spacings (0.05, 1, 1, 1), Qs = 0.5, amplitude 1,
coupling 1, 200 steps of StepWidth 1):
* the same seeds on both sides (paired ensembles): the ensemble means agree
  within three standard errors of the per-seed differences (or 1e-9
  relative, the resolution of trajectories that still track each other);
* different seeds on the two sides (independent ensembles of the same
  initial-condition distribution): the means agree within 3.5 combined
  standard errors.
Gauge fixing on both sides uses tolerance 1e-8, alpha 0.08 and at most 60
transformations; f(p) of one configuration is sqrt(AA EE) per momentum,
averaged over the momenta of each |p| bin.
"""
import numpy as np
import pytest

from oracle import gaugefix_oracle as gf

pytestmark = pytest.mark.gpu

EXT = (16, 16, 16)
SP = (0.05, 1.0, 1.0, 1.0)
STEPS = 200
NSEEDS = 12
GF = (1e-8, 0.08, 60)
BINS = (0.0, 0.5, 1.0, 1.5, 2.5, 4.0)


def _say(msg):
    print(msg, flush=True)


def _binned_f(pnorm, aa, ee):
    f = np.sqrt(np.maximum(aa * ee, 0.0))
    out = []
    for lo, hi in zip(BINS[:-1], BINS[1:]):
        sel = (pnorm > max(lo, 1e-12)) & (pnorm <= hi)
        out.append(f[sel].mean())
    return out


def _gpu_row(nrq, seed, pnorm):
    nrq.random_init(seed)
    g = nrq.GaugeConfiguration()
    g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
    g.Update(1.0, STEPS)
    o = g.observables()
    Ug, _ = g.download()
    g.CoulombGaugefixing(*GF)
    aa, ee = g.transverse_correlators()
    return [o["energy"] / np.prod(EXT), o["plaquette"]] + _binned_f(pnorm, aa, ee), Ug


def _oracle_row(oracle, lat, seed):
    U, E = oracle.box_init(seed, 0.5, 1.0, 1.0)
    oracle.leapfrog(U, E, 1.0, STEPS)
    row = [oracle.energy(U, E) / lat.vol, oracle.mean_plaquette(U)]
    W, _, _ = gf.coulomb_gauge_fix(lat, U, *GF)
    aa, ee = gf.transverse_correlators(lat, W, E)
    return row + _binned_f(lat.pnorm, aa, ee), U


NAMES = ["energy density", "plaquette"] + [f"f(|p| in ({lo}, {hi}])" for lo, hi in zip(BINS[:-1], BINS[1:])]


def _setup(nrq, oracle, lattice):
    lattice(EXT, SP)
    oracle.lattice_init(EXT, SP)
    return gf.Lattice(EXT, SP, oracle.generators())


def test_paired_ensembles_agree_after_decorrelation(nrq, oracle, lattice):
    lat = _setup(nrq, oracle, lattice)
    gpu, orc, link_diff = [], [], []
    for s in range(1, NSEEDS + 1):
        _say(f"paired ensemble: seed {s}, {STEPS} steps, gauge fixing and correlators")
        a, Ug = _gpu_row(nrq, s, lat.pnorm)
        b, U = _oracle_row(oracle, lat, s)
        gpu.append(a)
        orc.append(b)
        link_diff.append(float(np.max(np.abs(Ug - U))))
    gpu, orc = np.array(gpu), np.array(orc)
    _say(f"largest per-seed link difference after {STEPS} steps: {max(link_diff):.3e}")
    d = gpu - orc
    se = d.std(axis=0, ddof=1) / np.sqrt(NSEEDS)
    for k, name in enumerate(NAMES):
        m_g, m_o = gpu[:, k].mean(), orc[:, k].mean()
        _say(f"{name}: gpu {m_g:.10e} oracle {m_o:.10e} diff {m_g - m_o:.3e} se {se[k]:.3e}")
        assert abs(m_g - m_o) <= 3 * se[k] + 1e-9 * abs(m_o), name


def test_independent_ensembles_agree_within_statistical_errors(nrq, oracle, lattice):
    lat = _setup(nrq, oracle, lattice)
    gpu = np.array([_gpu_row(nrq, s, lat.pnorm)[0] for s in range(1, NSEEDS + 1)])
    orc = []
    for s in range(NSEEDS + 1, 2 * NSEEDS + 1):
        _say(f"independent ensemble: oracle seed {s}")
        orc.append(_oracle_row(oracle, lat, s)[0])
    orc = np.array(orc)
    for k, name in enumerate(NAMES):
        m_g, m_o = gpu[:, k].mean(), orc[:, k].mean()
        se = np.sqrt(gpu[:, k].var(ddof=1) / NSEEDS + orc[:, k].var(ddof=1) / NSEEDS)
        _say(f"{name}: gpu {m_g:.6e} +- {gpu[:, k].std(ddof=1) / np.sqrt(NSEEDS):.2e}, "
             f"oracle {m_o:.6e} +- {orc[:, k].std(ddof=1) / np.sqrt(NSEEDS):.2e}, {(m_g - m_o) / se:+.2f} sigma")
        assert abs(m_g - m_o) <= 3.5 * se, name
