"""Full-size workloads beside the single-configuration benchmark.

    python -m nrqcd_amd.workloads mixed      # 128^3 FP32 storage vs FP64, CN solvers at 128^3
    python -m nrqcd_amd.workloads ensemble   # 32 x 48^3 batched ensemble with 4 masses each
    python -m nrqcd_amd.workloads learned    # FGMRES + trained CNN vs GMRES at several step widths
    python -m nrqcd_amd.workloads gluon      # Coulomb gauge fixing + transverse correlators, f(p) at 64^3
    python -m nrqcd_amd.workloads spectrum   # mode-31 rows at 64^3 (CN step, 8 correlators, norms) and
                                             # spectral functions of 1024 correlators of 4097 points
    python -m nrqcd_amd.workloads full       # CN steps with c = (1,1,1,1) and EquilibriumInit at 64^3
    python -m nrqcd_amd.workloads cn         # CN steps/s per solver and stencil launch shape at 128^3,
                                             # FP32-stored and FP64 links (--shape nc64,nc32,tiling,xcd ...)
    python -m nrqcd_amd.workloads leapfrog   # one leapfrog Update call at 128^3 after the 1 GiB copy probe, for
                                             # counter runs (--precision 32|64, NRQ_LF_SHAPE pins a launch shape)

Each prints one JSON object.  Inputs are synthetic (the seeded box initial
condition of gaugeconfiguration_su3.f90:906-951); all timing uses HIP events
on the library's stream.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import (bandwidth_probe, cn_get_shape, cn_set_shape, GaugeConfiguration, leapfrog_timer_start, leapfrog_timer_stop, GetNorm2Momentum_G, NRQCDField, NRQ_PRECISION_FP32, NRQ_PRECISION_FP64, NRQ_SOLVER_CHEBYSHEV_FP64,
               NRQ_SOLVER_CHEBYSHEV_MIXED, NRQ_SOLVER_FGMRES_CNN, NRQ_SOLVER_GMRES_FP64, NRQ_SOLVER_GMRES_MIXED,
               PreconditionerNet, event_elapsed_ms, event_record, lattice_finalize, lattice_init, leapfrog_batch,
               nrqcd_finalize, nrqcd_init, nrqcd_set_preconditioner, nrqcd_set_solver, random_init, spectral_functions,
               synchronize, update_batch)

SPACINGS = (0.05, 1.0, 1.0, 1.0)
C1 = [1, 0, 0, 0]


def _timed(fn) -> float:
    synchronize()
    event_record(0)
    fn()
    event_record(1)
    return event_elapsed_ms(0, 1)


def mixed(size: int = 128, steps: int = 500, seed: int = 3, cn_steps: int = 5, log_every: int = 50) -> dict:
    """FP32 link/E storage with FP64 reunitarisation every 10 steps against the
    all-FP64 run from the same seed (observables in FP64 sums), then CN steps/s
    of every solver at this size on the FP64 configuration."""
    lattice_init((size,) * 3, SPACINGS)
    runs = {}
    for name, prec in (("fp64", NRQ_PRECISION_FP64), ("fp32", NRQ_PRECISION_FP32)):
        random_init(seed)
        g = GaugeConfiguration(prec)
        g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
        if prec == NRQ_PRECISION_FP32:
            g.SetReunitarisationInterval(10)
        shape = g.launch_shape()  # runs the autotuner (on scratch copies) outside the timed region
        log = [dict(step=0, **g.observables())]
        ms = 0.0
        for k in range(0, steps, log_every):
            n = min(log_every, steps - k)
            ms += _timed(lambda: g.Update(1.0, n))
            log.append(dict(step=k + n, **g.observables()))
        runs[name] = {"g": g, "log": log, "ms_per_step": ms / steps, "shape": shape}
    err = {}
    for key in ("energy", "gauss_max", "plaquette"):
        a = np.array([o[key] for o in runs["fp64"]["log"]])
        b = np.array([o[key] for o in runs["fp32"]["log"]])
        err[key] = {"max_abs": float(np.max(np.abs(a - b))),
                    "max_rel": float(np.max(np.abs(a - b) / np.maximum(np.abs(a), 1e-300)))}
    e = np.array([o["energy"] for o in runs["fp64"]["log"]])
    out = {
        "workload": f"{size}^3 box IC Qs=0.5 seed {seed}, {steps} leapfrog steps",
        "leapfrog_us_per_step": {k: 1e3 * v["ms_per_step"] for k, v in runs.items()},
        "link_updates_per_s": {k: 3 * size ** 3 / (1e-3 * v["ms_per_step"]) for k, v in runs.items()},
        "launch_shape": {k: v["shape"] for k, v in runs.items()},
        "fp32_vs_fp64_error": err,
        "fp64_energy_drift_rel": float((e[-1] - e[0]) / e[0]),
        "log": {k: v["log"] for k, v in runs.items()},
    }
    g64 = runs["fp64"]["g"]
    del runs
    nrqcd_init(2.0, C1, 1.0, 1e-10)
    cn = {}
    for name, kind in (("plain_fp64_gmres", NRQ_SOLVER_GMRES_FP64), ("mixed_gmres", NRQ_SOLVER_GMRES_MIXED),
                       ("fp64_chebyshev", NRQ_SOLVER_CHEBYSHEV_FP64), ("mixed_chebyshev", NRQ_SOLVER_CHEBYSHEV_MIXED)):
        nrqcd_set_solver(kind, 30, 10000)
        h = NRQCDField()
        h.InitSinglePoint(1, 1)
        h.Update(g64, 2.0, C1)
        stats = []
        ms = _timed(lambda: stats.extend(h.Update(g64, 2.0, C1) for _ in range(cn_steps)))
        cn[name] = {"cn_steps_per_s": cn_steps / (1e-3 * ms), "ms_per_step": ms / cn_steps,
                    "worst_true_relres": max(s.max_true_relres for s in stats),
                    "iterations_per_column_solve": sum(s.iterations for s in stats) / sum(s.solves for s in stats)}
        del h
    nrqcd_finalize()
    plain = cn["plain_fp64_gmres"]["cn_steps_per_s"]
    out["cn"] = cn
    out["cn_speedup_vs_plain"] = {k: v["cn_steps_per_s"] / plain for k, v in cn.items()}
    del g64
    lattice_finalize()
    return out


def ensemble(members: int = 32, size: int = 48, steps: int = 500, cn_steps: int = 10, precision: int = 32,
             first_seed: int = 100) -> dict:
    """`members` configurations (seeds first_seed.., Qs uniform in [0.3, 0.7]
    from a separate generator) advanced by one batched launch per pass in
    calls of 50 steps (one untimed call first), then
    CN steps of 4 heavy fields per configuration (Ma = 1, 2, 5, 10 at random
    sites) as one multi-right-hand-side solve per configuration."""
    lattice_init((size,) * 3, SPACINGS)
    vol = size ** 3
    rng = np.random.default_rng(1000)  # Qs and source sites, apart from the ranlxd stream of each seed
    qs = rng.uniform(0.3, 0.7, members)
    configs = []
    t0 = time.perf_counter()
    for k in range(members):
        random_init(first_seed + k)
        g = GaugeConfiguration(NRQ_PRECISION_FP32 if precision == 32 else NRQ_PRECISION_FP64)
        g.TransversePolarisedOccupiedInit_Box(float(qs[k]), 1.0, 1.0)
        if precision == 32:
            g.SetReunitarisationInterval(10)
        configs.append(g)
    init_s = time.perf_counter() - t0
    # calls of 50 steps; tuning and the capture of the call's graph outside the timed region
    call = 50 if steps % 50 == 0 else steps
    leapfrog_batch(configs, 1.0, call)

    def run():
        for _ in range(steps // call):
            leapfrog_batch(configs, 1.0, call)

    ms = _timed(run)
    e = [c.GetEnergy() for c in configs]
    out = {
        "workload": f"{members} x {size}^3 ensemble, FP{precision} storage, seeds {first_seed}..{first_seed + members - 1}",
        "initial_conditions_s": init_s,
        "leapfrog_ms_per_step": ms / steps,
        "link_updates_per_s": 3 * vol * members * steps / (1e-3 * ms),
        "energies": e,
    }
    masses = [1.0, 2.0, 5.0, 10.0]
    nrqcd_init(2.0, C1, 1.0, 1e-10)
    nrqcd_set_solver(NRQ_SOLVER_CHEBYSHEV_MIXED, 30, 10000)
    fields = []
    for k in range(members):
        sites = rng.integers(1, vol + 1, size=(4, 2))
        fs = []
        for (iq, ia) in sites:
            h = NRQCDField()
            h.InitSinglePoint(int(iq), int(ia))
            fs.append(h)
        fields.append(fs)
    for k in range(members):
        update_batch(fields[k], configs[k], masses, C1)  # workspace
    stats = []

    def cn_round():
        for _ in range(cn_steps):
            for k in range(members):
                stats.append(update_batch(fields[k], configs[k], masses, C1))

    ms = _timed(cn_round)
    out["cn"] = {"solver": "FP32 Chebyshev corrections, FP64 outer residual 1e-10",
                 "configuration_cn_steps_per_s": members * cn_steps / (1e-3 * ms),
                 "column_solves_per_s": sum(s.solves for s in stats) / (1e-3 * ms),
                 "worst_true_relres": max(s.max_true_relres for s in stats),
                 "norms_member0": [f.norms() for f in fields[0]]}
    del fields
    nrqcd_finalize()
    del configs
    lattice_finalize()
    return out


def learned(size: int = 32, stepwidths=(1.0, 10.0, 40.0), seed: int = 7, evolve: int = 40, solves: int = 6,
            epochs: int = 30, lr: float = 1e-3, cn_steps: int = 3) -> dict:
    """Train the preconditioner network on pairs from plain GMRES solves at
    Ma = 1, 2, 5 and compare FGMRES(20) + network with GMRES(20) on held-out
    masses 3 and 8 (point sources, zero initial guess, FP64 true residual
    1e-10), for each step width on an evolved configuration."""
    lattice_init((size,) * 3, SPACINGS)
    nrqcd_init(2.0, C1, 1.0, 1e-10)
    random_init(seed)
    g = GaugeConfiguration()
    g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
    g.Update(1.0, evolve)
    rows = []
    for sw in stepwidths:
        net = PreconditionerNet(3)
        t0 = time.perf_counter()
        for m in (1.0, 2.0, 5.0):
            net.harvest(g, m, solves, seed=int(10 * m), StepWidth=sw)
        loss0 = net.train(1, lr, seed=1)
        loss1 = net.train(epochs - 1, lr, seed=2)
        train_s = time.perf_counter() - t0
        nrqcd_set_preconditioner(net)
        for m in (3.0, 8.0):
            row = {"stepwidth": sw, "mass": m, "beta_kmax": sw * SPACINGS[0] / 2 / m * 12,
                   "pairs": net.dataset_size(), "loss_first_epoch": loss0, "loss_last_epoch": loss1,
                   "training_s": train_s}
            for name, kind in (("gmres", NRQ_SOLVER_GMRES_FP64), ("fgmres_cnn", NRQ_SOLVER_FGMRES_CNN)):
                nrqcd_set_solver(kind, 20, 10000)
                h = NRQCDField()
                h.InitSinglePoint(100, 900)
                h.Update(g, m, C1, sw)
                stats = []
                ms = _timed(lambda: stats.extend(h.Update(g, m, C1, sw) for _ in range(cn_steps)))
                row[name] = {"iterations_per_column_solve": sum(s.iterations for s in stats) / sum(s.solves for s in stats),
                             "ms_per_cn_step": ms / cn_steps,
                             "worst_true_relres": max(s.max_true_relres for s in stats)}
                del h
            rows.append(row)
        nrqcd_set_preconditioner(None)
        del net
    nrqcd_finalize()
    del g
    lattice_finalize()
    return {"workload": f"{size}^3 evolved box IC (seed {seed}, {evolve} steps), train at Ma 1,2,5, test at 3,8",
            "rows": rows}


def gluon(size: int = 64, steps: int = 20, seed: int = 2, iterations: int = 200, tol: float = 1e-8,
          alpha: float = 0.08, precision: int = 64) -> dict:
    """Gluon distribution of the evolved box IC as the reference's mode 1
    computes it (main.f90:3020-3190): Coulomb gauge fixing of a copy (at most
    `iterations` transformations), transverse AA and EE correlators,
    f(p) = sqrt(AA EE); f is also averaged in |p| bins of width 0.25."""
    lattice_init((size,) * 3, SPACINGS)
    random_init(seed)
    g = GaugeConfiguration(NRQ_PRECISION_FP32 if precision == 32 else NRQ_PRECISION_FP64)
    g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
    g.Update(1.0, steps)
    warm = g.copy()
    warm.CoulombGaugefixing(tol, alpha, 2)  # hipFFT plan and kernels outside the timed region
    warm.transverse_correlators()
    del warm
    fixed = g.copy()
    done = [0, 0.0]
    ms_fix = _timed(lambda: done.__setitem__(slice(0, 2), fixed.CoulombGaugefixing(tol, alpha, iterations)))
    corr = {}
    ms_corr = _timed(lambda: corr.setdefault("aaee", fixed.transverse_correlators()))
    aa, ee = corr["aaee"]
    f = np.sqrt(np.maximum(aa * ee, 0.0))
    p = np.array([GetNorm2Momentum_G(k) for k in range(1, size ** 3 + 1)])
    edges = np.arange(0.0, p.max() + 0.25, 0.25)
    idx = np.digitize(p, edges)
    bins = [{"p": float(p[idx == k].mean()), "f": float(f[idx == k].mean()), "modes": int((idx == k).sum())}
            for k in np.unique(idx) if k > 0 and (idx == k).sum() > 0 and p[idx == k].mean() > 0]
    out = {
        "workload": f"{size}^3 box IC Qs=0.5 seed {seed}, {steps} leapfrog steps, FP{precision} links, "
                    f"gauge fixing tol {tol} alpha {alpha} at most {iterations} transformations",
        "transformations": done[0], "final_max_deviation": done[1],
        "gaugefix_ms": ms_fix, "gaugefix_us_per_iteration": 1e3 * ms_fix / max(done[0], 1),
        "correlators_ms": ms_corr, "f_binned": bins,
    }
    del fixed, g
    lattice_finalize()
    return out


def spectrum(size: int = 64, steps: int = 8, seed: int = 2, reps: int = 10, batch: int = 1024,
             points: int = 4097) -> dict:
    """The measurement loop of the reference's mode 31 (main.f90:586-619) at
    `size`^3: per row a CN step and a leapfrog step, then the eight
    zero-momentum correlators and the two norms (timed against the same rows
    without them); and mode 4's spectral
    function (main.f90:1865-2100) for `batch` correlators of `points` time
    points at once (host arrays in and out, so the time includes the PCIe
    copies)."""
    lattice_init((size,) * 3, SPACINGS)
    random_init(seed)
    nrqcd_init(2.0, C1, 1.0, 1e-10)
    nrqcd_set_solver(NRQ_SOLVER_CHEBYSHEV_MIXED, 30, 10000)
    g = GaugeConfiguration()
    g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
    h = NRQCDField()
    h.InitSinglePoint(1, 1)
    for _ in range(steps):
        h.Update(g, 2.0, C1)
        g.Update(1.0, 1)
    h.correlators()

    def rows(observe: bool):
        for _ in range(reps):
            h.Update(g, 2.0, C1)
            g.Update(1.0, 1)
            if observe:  # what the driver writes per row: 8 correlators, 2 norms
                h.correlators()
                h.correlators(True)
                h.norms()

    ms_row = _timed(lambda: rows(True)) / reps
    ms_evolve = _timed(lambda: rows(False)) / reps
    c = h.correlators()
    cg = h.correlators(True)
    del h, g
    nrqcd_finalize()
    lattice_finalize()
    rng = np.random.default_rng(seed)
    t = np.arange(points) * SPACINGS[0]
    e = rng.uniform(0.5, 4.0, size=(batch, 1))
    sig = 1j * np.exp(-1j * e * t[None, :]) * np.exp(-0.01 * t[None, :])
    spectral_functions(sig, SPACINGS[0])
    t0 = time.perf_counter()
    spec = spectral_functions(sig, SPACINGS[0])
    ms_spec = 1e3 * (time.perf_counter() - t0)
    peak = np.argmax(spec, axis=1) * 2 * np.pi / (SPACINGS[0] * spec.shape[1])
    return {
        "workload": f"{size}^3 box IC Qs=0.5 seed {seed}, point source, {steps} CN + leapfrog steps, "
                    f"mixed Chebyshev CN; {batch} correlators of {points} points",
        "row_ms": ms_row, "cn_and_leapfrog_ms": ms_evolve, "observables_ms": ms_row - ms_evolve,
        "O1_1S0": [c["O1_1S0"].real, c["O1_1S0"].imag], "O1_1S0_gconjg": [cg["O1_1S0"].real, cg["O1_1S0"].imag],
        "spectra_ms_with_pcie": ms_spec,
        "spectra_peak_max_error": float(np.max(np.abs(peak - e[:, 0]))),
    }


def full(size: int = 64, steps: int = 3, seed: int = 2, beta: float = 16.0, nequilibrium: int = 10) -> dict:
    """The CN step with the full Hamiltonian (c = (1, 1, 1, 1), the reference
    templates' coefficients; nrqcd.f90:1175-1387) against the kinetic-only
    step with the same FP64 GMRES, and one pass of EquilibriumInit
    (gaugeconfiguration_su3.f90:380-814) at `size`^3."""
    lattice_init((size,) * 3, SPACINGS)
    random_init(seed)
    nrqcd_init(2.0, C1, 1.0, 1e-10)
    nrqcd_set_solver(NRQ_SOLVER_GMRES_FP64, 30, 10000)
    g = GaugeConfiguration()
    g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
    g.Update(1.0, 20)
    out = {"workload": f"{size}^3 box IC Qs=0.5 seed {seed}, 20 leapfrog steps, point source, Ma=2, FP64 GMRES(30) "
                       f"to 1e-10; EquilibriumInit beta={beta}, one pass, {nequilibrium} leapfrog steps"}
    for name, c in (("kinetic", [1, 0, 0, 0]), ("full", [1, 1, 1, 1])):
        h = NRQCDField()
        h.InitSinglePoint(1, 1)
        h.Update(g, 2.0, c)  # warm-up (workspace, first step)
        stats = []
        ms = _timed(lambda: [stats.append(h.Update(g, 2.0, c)) for _ in range(steps)]) / steps
        out[name] = {"ms_per_cn_step": ms, "iterations_per_column_solve": stats[-1].iterations / max(stats[-1].solves, 1),
                     "worst_true_relres": max(s.max_true_relres for s in stats)}
        del h
    out["full_over_kinetic"] = out["full"]["ms_per_cn_step"] / out["kinetic"]["ms_per_cn_step"]
    del g
    e = GaugeConfiguration()
    rec = {}
    ms_eq = _timed(lambda: rec.setdefault("r", e.EquilibriumInit(beta, 1, nequilibrium)))
    en, dev, nproj = rec["r"]
    out["equilibrium"] = {"ms": ms_eq, "projection_sweeps": int(nproj[0]), "ms_per_sweep_upper_bound": ms_eq / max(int(nproj[0]), 1),
                          "energy": float(en[0]), "gauss_deviation": float(dev[0])}
    del e
    nrqcd_finalize()
    lattice_finalize()
    return out


def cn(size: int = 128, evolve: int = 50, seed: int = 3, cn_steps: int = 5, shapes=None,
       solvers=("mixed_chebyshev", "plain_fp64_gmres")) -> dict:
    """CN steps of point sources (Ma = 2, 12 fields, FP64 true relative
    residual 1e-10) on the seeded box configuration evolved `evolve` steps,
    stored in FP32 (reunitarised every 10 steps) and in FP64, for each
    stencil launch shape (nc64, nc32, tiling, xcd) in turn; every shape twice,
    interleaved, so that both passes see the same machine state."""
    kinds = {"mixed_chebyshev": NRQ_SOLVER_CHEBYSHEV_MIXED, "plain_fp64_gmres": NRQ_SOLVER_GMRES_FP64,
             "mixed_gmres": NRQ_SOLVER_GMRES_MIXED, "fp64_chebyshev": NRQ_SOLVER_CHEBYSHEV_FP64}
    lattice_init((size,) * 3, SPACINGS)
    confs = {}
    for name, prec in (("fp32_links", NRQ_PRECISION_FP32), ("fp64_links", NRQ_PRECISION_FP64)):
        random_init(seed)
        g = GaugeConfiguration(prec)
        g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
        if prec == NRQ_PRECISION_FP32:
            g.SetReunitarisationInterval(10)
        g.Update(1.0, evolve)
        confs[name] = g
    nrqcd_init(2.0, C1, 1.0, 1e-10)
    default = cn_get_shape()
    shapes = [tuple(sh) for sh in shapes] if shapes else [default]
    out = {"size": size, "evolve": evolve, "cn_steps": cn_steps, "default_shape": list(default), "runs": []}
    try:
        for rep in range(2):
            for sh in shapes:
                cn_set_shape(*sh)
                for conf, g in confs.items():
                    for sv in solvers:
                        if sv == "plain_fp64_gmres" and conf == "fp32_links":
                            continue
                        nrqcd_set_solver(kinds[sv], 30, 10000)
                        h = NRQCDField()
                        h.InitSinglePoint(1, 1)
                        h.Update(g, 2.0, C1)
                        n = cn_steps if "chebyshev" in sv else max(1, cn_steps // 2)
                        its = []
                        ms = _timed(lambda: its.extend(h.Update(g, 2.0, C1).iterations for _ in range(n)))
                        out["runs"].append({"rep": rep, "shape": list(sh), "links": conf, "solver": sv,
                                            "ms_per_step": ms / n, "iterations_per_step": sum(its) / n})
                        del h
    finally:
        cn_set_shape(*default)
        nrqcd_finalize()
        confs.clear()
        lattice_finalize()
    return out


def leapfrog(size: int = 128, precision: int = 32, steps: int = 10, seed: int = 3) -> dict:
    """The seeded box configuration (Qs = 0.5) at this storage precision, the
    launch shape chosen (or pinned by NRQ_LF_SHAPE), the 1 GiB copy kernel
    (k_stream_copy, a calibration point for HBM counters) and one
    `steps`-step Update call (steps - 1 fused launches) timed per fused launch."""
    lattice_init((size,) * 3, SPACINGS)
    try:
        random_init(seed)
        g = GaugeConfiguration(NRQ_PRECISION_FP32 if precision == 32 else NRQ_PRECISION_FP64)
        g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
        shape = g.launch_shape()
        copy_gbs = bandwidth_probe(1 << 30, 2)
        synchronize()
        leapfrog_timer_start(steps + 1)
        g.Update(1.0, steps)
        t = leapfrog_timer_stop()
        del g
    finally:
        lattice_finalize()
    return {"size": size, "precision": precision, "steps": steps, "launch_shape": shape, "copy_gbs": copy_gbs,
            "fused_launches": t["launches"], "fused_us": 1e3 * t["ms"] / max(t["launches"], 1)}


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("workload", choices=("mixed", "ensemble", "learned", "gluon", "spectrum", "full", "cn", "leapfrog"))
    ap.add_argument("--size", type=int)
    ap.add_argument("--steps", type=int)
    ap.add_argument("--members", type=int, default=32)
    ap.add_argument("--cn-steps", type=int)
    ap.add_argument("--shape", action="append", help="cn: stencil launch shape nc64,nc32,tiling,xcd (repeatable)")
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64), help="leapfrog: storage precision")
    a = ap.parse_args(argv)
    kw = {k: v for k, v in (("size", a.size), ("steps", a.steps), ("cn_steps", a.cn_steps)) if v is not None}
    if a.workload == "mixed":
        out = mixed(**kw)
    elif a.workload == "learned":
        out = learned(**{k: v for k, v in kw.items() if k in ("size", "cn_steps")})
    elif a.workload == "gluon":
        out = gluon(**{k: v for k, v in kw.items() if k in ("size", "steps")})
    elif a.workload == "full":
        out = full(**{k: v for k, v in kw.items() if k in ("size", "steps")})
    elif a.workload == "leapfrog":
        out = leapfrog(precision=a.precision, **{k: v for k, v in kw.items() if k in ("size", "steps")})
    elif a.workload == "cn":
        shapes = [tuple(int(x) for x in sh.split(",")) for sh in a.shape] if a.shape else None
        out = cn(shapes=shapes, **{k: v for k, v in kw.items() if k in ("size", "cn_steps")})
    elif a.workload == "spectrum":
        out = spectrum(**{k: v for k, v in kw.items() if k in ("size", "steps")})
    else:
        out = ensemble(members=a.members, **kw)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    sys.exit(main())
