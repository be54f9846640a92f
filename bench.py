"""Throughput of the real-time SU(3) leapfrog + NRQCD Crank-Nicolson hot path
on the 128^3 mixed-precision workload.

Workload (one GPU): 128^3 lattice, spacings (a0, a1, a2, a3) = (0.05, 1, 1, 1),
transverse-polarised box initial condition with Qs = 0.5, amplitude 1,
coupling 1, ranlxd seed 3.

* Leapfrog (the headline `value`): FP32 link and E storage, reunitarised in
  FP64 every 10 steps, energy / Gauss law / plaquette summed in FP64 every
  LOG_EVERY steps inside the timed region; `--steps` steps.  The all-FP64 run
  from the same seed is timed beside it, and the FP32 run's error per
  observable against it is reported.
* CN at 128^3 on the FP64 configuration: heavy quark and antiquark point
  sources at lattice index 1, Ma = 2, c = (1, 0, 0, 0), every column solve to
  an FP64 true relative residual of 1e-10.  Solvers: plain FP64 GMRES(30)
  (PETSc's algorithm; Jacobi preconditioning is a constant scaling of this
  operator, so plain GMRES is Jacobi-GMRES), FP32 GMRES in FP64 iterative
  refinement, FP64 Chebyshev, FP32 Chebyshev in FP64 refinement, and the
  learned preconditioner in FGMRES(20) (FP64, and FP32 inside FP64
  refinement).
* Learned preconditioner: the seed-3 configuration evolved to t = 10/Qs
  (400 steps), pairs harvested from plain GMRES solves at Ma = 1, 2, 5,
  the network trained on the GPU, then FGMRES(20) with it against
  GMRES(20) at the held-out Ma = 3 and 8 (iterations and CN steps/s).

All GPU times are HIP events on the library stream with the inputs resident
in HBM.  The roofline is that of the fused leapfrog kernel (FP32 storage, 624
algorithmic bytes per site per launch) over its own event-timed launches.

Prints one JSON line.  Options: --gpus N (only 1), --steps K, --warmup W,
--size L, --cn-steps C, --no-cpu (skip the CPU baseline), --no-learned.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "link-updates/s & CN steps/s at FP64 rel.res 1e-10: learned+mixed vs plain"
SPACINGS = (0.05, 1.0, 1.0, 1.0)
SEED = 3
C1 = [1, 0, 0, 0]
LOG_EVERY = 50        # steps between logged observables (energy, Gauss law, plaquette)
REUNIT_EVERY = 10     # FP32 storage: FP64 reunitarisation interval
HBM_PEAK_GBS = 8000.0
EPOCHS = 10           # training passes of the learned preconditioner
# Algorithmic HBM bytes per site of one fused leapfrog launch (E update of one
# step with the link update of the next): U (3 x 9 complex) and E (3 x 8
# reals) read once and written once.
FUSED_BYTES = {64: 1248.0, 32: 624.0}
# Algorithmic HBM bytes per site of one mixed-Chebyshev CN step with 12 live
# fields when every column takes the planned 6 Chebyshev steps (two FP32
# corrections of 4 and 2 steps), FP64 links: right-hand side with y, x0 (FP64)
# and K psi (FP32) 2304; FP32 copy of two link rows 432; K pass 720; Chebyshev
# launches 720, 1008, 2160 (adds the correction to x0 in FP64) and 1584 (adds
# the second to psi); two FP64 true-residual passes 2 x 1728.  Every pass reads
# two rows of each link (288 B per site in FP64, 144 B in FP32).
CN_MIXED_CHEB_BYTES = 2304.0 + 432 + 720 + 720 + 1008 + 2160 + 1584 + 2 * 1728
CN_MIXED_CHEB_ITS = 6.0
TRAFFIC_FILE = ROOT / "profiles" / "leapfrog_fused_128cubed_traffic.txt"
# one network apply: 2 flop per multiply-add, 27 taps, 18 -> 32 -> 32 -> 32 -> 18 channels, per site
CNN_FLOP_PER_SITE = 2 * 27 * (18 * 32 + 32 * 32 + 32 * 32 + 32 * 18)
BF16_MFMA_PEAK_TFLOPS = 2500.0
TILINGS = ("consecutive sites", "z-sweep tiles of x-rows", "z-sweep 16-wide tiles", "LDS tiles, 16-plane runs",
           "LDS tiles, 8-plane runs")


def _shape_key(precision: int, shape: dict) -> str:
    return (f"fp{precision} bs={shape['threads_per_block']} xcd={shape['xcd_remap']} "
            f"tiling={TILINGS.index(shape['tiling'])} streaming={int(shape['streaming'])} "
            f"two_row={int(shape['two_row_neighbours'])}")


def _profiled_traffic(size: int, precision: int, shape: dict):
    """HBM bytes per fused launch of this launch shape as rocprofv3 measured
    them (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel) in the
    committed profile, or None when that shape was not measured."""
    if size != 128 or not TRAFFIC_FILE.exists():
        return None
    key = f"traffic_bytes_per_launch[{_shape_key(precision, shape)}]="
    for line in TRAFFIC_FILE.read_text().splitlines():
        if line.startswith(key):
            return float(line[len(key):])
    return None


def _leapfrog_leg(nrq, size: int, precision: int, steps: int, warmup: int) -> dict:
    """One configuration of the workload at this storage precision: warm-up
    (launch-shape tuning included), then `steps` timed steps with the
    observables logged every LOG_EVERY steps on the device."""
    V = size ** 3
    nrq.random_init(SEED)
    g = nrq.GaugeConfiguration(nrq.NRQ_PRECISION_FP64 if precision == 64 else nrq.NRQ_PRECISION_FP32)
    g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
    if precision == 32:
        g.SetReunitarisationInterval(REUNIT_EVERY)
    shape = g.launch_shape()  # the autotuner runs here, on scratch copies
    obs0 = g.observables()
    if warmup:
        w = g.copy()
        w.Update(1.0, warmup)
        del w
    g.log_observables()
    nrq.synchronize()
    nrq.leapfrog_timer_start(steps + 1)
    nrq.event_record(0)
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        n = min(LOG_EVERY, steps - done)
        g.Update(1.0, n)
        g.log_observables()
        done += n
    nrq.event_record(1)
    tm = nrq.leapfrog_timer_stop()  # synchronises
    wall = time.perf_counter() - t0
    ms = nrq.event_elapsed_ms(0, 1)
    log = g.read_log()
    fused_us = 1e3 * tm["ms"] / max(tm["launches"], 1)
    achieved = FUSED_BYTES[precision] * V / (fused_us * 1e-6) / 1e9
    return {
        "g": g,
        "log": log,
        "out": {
            "link_updates_per_s": 3 * V * steps / (ms * 1e-3),
            "ms_per_step": ms / steps,
            "wall_s": wall,
            "launch_shape": shape,
            "fused_kernel_us": fused_us,
            "fused_launches_timed": tm["launches"],
            "fused_gbs_algorithmic": achieved,
            "fused_frac_of_peak": achieved / HBM_PEAK_GBS,
            "energy_drift_rel": float((log[-1][0] - log[0][0]) / log[0][0]),
            "gauss_max_start_end": [float(log[0][2]), float(log[-1][2])],
            "energy_start_ic": obs0["energy"],
        },
    }


def _cn_leg(nrq, g, kind, restart: int, mass: float, cn_steps: int, source=(1, 1)) -> dict:
    nrq.nrqcd_set_solver(kind, restart, 10000)
    h = nrq.NRQCDField()
    h.InitSinglePoint(*source)
    h.Update(g, mass, C1)  # warm-up (workspace allocation, first step)
    nrq.synchronize()
    its = solves = 0
    worst = 0.0
    nrq.event_record(2)
    for _ in range(cn_steps):
        st = h.Update(g, mass, C1)
        its += st.iterations
        solves += st.solves
        worst = max(worst, st.max_true_relres)
    nrq.event_record(3)
    ms = nrq.event_elapsed_ms(2, 3)
    nq, na = h.norms()
    del h
    return {"cn_steps_per_s": cn_steps / (ms * 1e-3), "ms_per_step": ms / cn_steps,
            "column_solves_per_s": solves / (ms * 1e-3), "iterations_per_column_solve": its / max(solves, 1),
            "worst_true_relres": worst, "norms": [nq, na]}


def _train(nrq, g_ic):
    """The learned-preconditioner data: the seed-3 initial condition evolved
    to t = 10/Qs (400 steps of 0.05 a), pairs from one plain GMRES solve of
    12 random fields at each of Ma = 1, 2, 5 (36 whole-lattice pairs, each
    512 16^3 patches), EPOCHS epochs of Adam on the GPU."""
    g = g_ic.copy()
    g.Update(1.0, 400)
    net = nrq.PreconditionerNet(3)
    t0 = time.perf_counter()
    for m in (1.0, 2.0, 5.0):
        net.harvest(g, m, 1, seed=int(10 * m))
    loss0 = net.train(1, 1e-3, seed=1)
    loss1 = net.train(EPOCHS - 1, 1e-3, seed=2)
    nrq.synchronize()
    info = {"configuration": f"seed {SEED} evolved 400 steps (t = 10/Qs)", "training_pairs": net.dataset_size(),
            "patches_16cubed": net.dataset_size() * (nrq.lattice_size() // 4096), "epochs": EPOCHS,
            "loss_first_epoch": loss0, "loss_last_epoch": loss1, "harvest_and_training_s": time.perf_counter() - t0}
    info["network_apply_ms"] = net.time_forward(20)
    return net, g, info


def _held_out(nrq, g, net, cn_steps: int) -> dict:
    """FGMRES(20) with the trained network (FP64, and FP32 in FP64
    refinement) against GMRES(20), mixed GMRES(20) and FGMRES(20) with the
    untrained (zero) network, i.e. the second-order Neumann polynomial the
    network corrects, at the held-out Ma = 3 and 8 on the evolved
    configuration (point sources, FP64 true residual 1e-10)."""
    zero = nrq.PreconditionerNet(3)
    zero.weights = np.zeros(nrq.NRQ_CNN_NWEIGHTS, np.float32)
    out = {}
    for m in (3.0, 8.0):
        row = {}
        for name, kind, pre in (("jacobi_gmres", nrq.NRQ_SOLVER_GMRES_FP64, None),
                                ("polynomial_fgmres", nrq.NRQ_SOLVER_FGMRES_CNN, zero),
                                ("learned_fgmres", nrq.NRQ_SOLVER_FGMRES_CNN, net),
                                ("mixed_gmres", nrq.NRQ_SOLVER_GMRES_MIXED, None),
                                ("learned_mixed_fgmres", nrq.NRQ_SOLVER_FGMRES_CNN_MIXED, net)):
            nrq.nrqcd_set_preconditioner(pre)
            before = nrq.cn_fallbacks()
            row[name] = _cn_leg(nrq, g, kind, 20, m, cn_steps)
            row[name]["fallbacks_to_gmres"] = nrq.cn_fallbacks() - before
        base = row["jacobi_gmres"]
        row["learned_vs_jacobi_gmres"] = {
            "cn_steps_per_s": row["learned_fgmres"]["cn_steps_per_s"] / base["cn_steps_per_s"],
            "iterations": row["learned_fgmres"]["iterations_per_column_solve"] / base["iterations_per_column_solve"]}
        out[str(m)] = row
    nrq.nrqcd_set_preconditioner(net)
    del zero
    return out


def _cpu_baseline(U, E, size: int, seconds: float) -> dict:
    """The oracle (OpenMP FP64 C restatement) on bounded samples of the same
    workload: leapfrog steps of the 128^3 seed-3 initial condition (the
    observables every LOG_EVERY steps are not reached in the sample), and CN
    steps/s of the 16^3 CPU configuration (seed 1, Ma = 2, GMRES(30) to
    relative residual 1e-12)."""
    from oracle.pyoracle import Oracle

    orc = Oracle()
    orc.lattice_init((size,) * 3, SPACINGS)
    t0 = time.perf_counter()
    orc.leapfrog(U, E, 1.0, 1)
    one = time.perf_counter() - t0
    steps = max(1, int(0.5 * seconds / max(one, 1e-6)))
    t0 = time.perf_counter()
    orc.leapfrog(U, E, 1.0, steps)
    dt = time.perf_counter() - t0
    lu = 3 * size ** 3 * steps / dt
    orc.lattice_init((16, 16, 16), SPACINGS)
    U16, E16 = orc.box_init(1, 0.5, 1.0, 1.0)
    Q, A = orc.point_prop(1), orc.point_prop(1)
    t0 = time.perf_counter()
    orc.cn_update(U16, Q, A, 2.0, C1, 1.0, 1e-12)
    one = time.perf_counter() - t0
    cn_n = max(1, int(0.5 * seconds / max(one, 1e-6)))
    t0 = time.perf_counter()
    for _ in range(cn_n):
        orc.cn_update(U16, Q, A, 2.0, C1, 1.0, 1e-12)
    cn_dt = time.perf_counter() - t0
    return {
        "value": lu,
        "unit": "link-updates/s",
        "cores": orc.num_threads(),
        "kind": "port",
        "sample": f"oracle FP64 leapfrog of the {size}^3 seed-{SEED} initial condition, {steps} steps after one "
                  f"warm-up step ({dt:.1f} s)",
        "cn": {"value": cn_n / cn_dt, "unit": "CN steps/s", "cores": orc.num_threads(), "kind": "port",
               "sample": f"oracle CN steps (GMRES(30), classical Gram-Schmidt, relative residual 1e-12) of a 16^3 "
                         f"seed-1 point source at Ma = 2, {cn_n} steps after one warm-up step ({cn_dt:.1f} s)"},
    }


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--cn-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for {args.gpus}); the lattice is not "
              "sharded across GPUs", file=sys.stderr)
        sys.exit(2)

    import nrqcd_amd as nrq

    L = args.size
    V = L ** 3
    nrq.lattice_init((L, L, L), SPACINGS)

    # ---- leapfrog: FP32 storage (headline) and the FP64 run beside it --------
    legs, logs, gs = {}, {}, {}
    for prec in (32, 64):
        r = _leapfrog_leg(nrq, L, prec, args.steps, args.warmup)
        legs[f"fp{prec}"] = r["out"]
        logs[prec] = r["log"]
        gs[prec] = r["g"]
    g32, g64 = gs[32], gs[64]
    del gs
    err = {}
    for k, key in enumerate(("energy", "gauss_ref", "gauss_max", "plaquette")):
        a, b = logs[64][:, k], logs[32][:, k]
        err[key] = {"max_abs": float(np.max(np.abs(a - b))),
                    "max_rel": float(np.max(np.abs(a - b) / np.maximum(np.abs(a), 1e-300)))}
    lf = legs["fp32"]

    # ---- CN at 128^3 ---------------------------------------------------------
    nrq.random_init(SEED)
    g_ic = nrq.GaugeConfiguration(nrq.NRQ_PRECISION_FP64)
    g_ic.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
    nrq.nrqcd_init(2.0, C1, 1.0, 1e-10)
    learned = net = g_ev = None
    if not args.no_learned:
        net, g_ev, learned = _train(nrq, g_ic)
        nrq.nrqcd_set_preconditioner(net)
    solvers = [("plain_fp64_gmres", nrq.NRQ_SOLVER_GMRES_FP64), ("mixed_gmres", nrq.NRQ_SOLVER_GMRES_MIXED),
               ("fp64_chebyshev", nrq.NRQ_SOLVER_CHEBYSHEV_FP64), ("mixed_chebyshev", nrq.NRQ_SOLVER_CHEBYSHEV_MIXED)]
    if net is not None:
        solvers += [("learned_fgmres", nrq.NRQ_SOLVER_FGMRES_CNN),
                    ("learned_mixed_fgmres", nrq.NRQ_SOLVER_FGMRES_CNN_MIXED)]
    cn = {name: _cn_leg(nrq, g64, kind, 30, 2.0, args.cn_steps) for name, kind in solvers}
    # the same mixed solvers on the FP32-stored configuration (FP32 links in
    # every pass; the FP64 outer residual uses them promoted)
    cn_fp32_links = {name: _cn_leg(nrq, g32, kind, 30, 2.0, args.cn_steps) for name, kind in solvers
                     if "mixed" in name}
    if net is not None:
        learned["held_out"] = _held_out(nrq, g_ev, net, args.cn_steps)
        nrq.nrqcd_set_preconditioner(None)
        del net, g_ev
    nrq.nrqcd_finalize()
    copy_gbs = nrq.bandwidth_probe(1 << 30, 20)
    plain = cn["plain_fp64_gmres"]["cn_steps_per_s"]
    mc = cn["mixed_chebyshev"]
    cn_roof = None
    if mc["iterations_per_column_solve"] == CN_MIXED_CHEB_ITS:
        gbs = CN_MIXED_CHEB_BYTES * V / (mc["ms_per_step"] * 1e-3) / 1e9
        cn_roof = {"solver": "mixed_chebyshev", "bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": gbs / HBM_PEAK_GBS, "bytes_per_step": CN_MIXED_CHEB_BYTES * V}

    cnn_roof = None
    if learned is not None:
        tf = CNN_FLOP_PER_SITE * V / (learned["network_apply_ms"] * 1e-3) / 1e12
        cnn_roof = {"kernel": "k_conv x 4 layers (one network apply, one colour field and its features)", "bound": "mfma",
                    "achieved": tf, "peak": BF16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tf / BF16_MFMA_PEAK_TFLOPS,
                    "flop_per_apply": CNN_FLOP_PER_SITE * V, "ms_per_apply": learned["network_apply_ms"]}

    cpu = None
    if not args.no_cpu:
        U, E = g_ic.download()
        del g_ic
        g_ic = None
        cpu = _cpu_baseline(U, E, L, args.cpu_seconds)
        del U, E

    shape = lf["launch_shape"]
    out = {
        "metric": METRIC,
        "value": lf["link_updates_per_s"],
        "unit": "link-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": lf["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{L}^3 SU(3) leapfrog, FP32 link/E storage with FP64 reunitarisation every "
                        f"{REUNIT_EVERY} steps and FP64 sums, box IC Qs=0.5 seed {SEED}, {args.steps} steps; "
                        f"NRQCD CN point source Ma=2 at {L}^3 to FP64 true rel.res 1e-10",
            "lattice": [L, L, L],
            "spacings": list(SPACINGS),
            "observables_every": LOG_EVERY,
            "cn_solvers": "GMRES(30) FP64 (plain), FP32-GMRES/FP64-IR, Chebyshev FP64, FP32-Chebyshev/FP64-IR, "
                          "learned FGMRES(30) FP64 and FP32/FP64-IR; learned FGMRES(20) vs GMRES(20) at Ma 3, 8",
            "mass": 2.0,
        },
        "leapfrog": legs,
        "fp32_vs_fp64_error": err,
        "cn_steps_per_s": mc["cn_steps_per_s"],
        "cn_steps_per_s_plain": plain,
        "cn": cn,
        "cn_speedup_vs_plain": {k: v["cn_steps_per_s"] / plain for k, v in cn.items()},
        "cn_on_fp32_stored_links": cn_fp32_links,
        "cn_on_fp32_stored_links_speedup_vs_plain": {k: v["cn_steps_per_s"] / plain for k, v in cn_fp32_links.items()},
        "cn_mixed_vs_fp64_chebyshev": mc["cn_steps_per_s"] / cn["fp64_chebyshev"]["cn_steps_per_s"],
        "cn_roofline": cn_roof,
        "cnn_roofline": cnn_roof,
        "learned": learned,
        "hbm_copy_probe_gbs": copy_gbs,
        "roofline": {
            "bound": "hbm",
            "achieved": lf["fused_gbs_algorithmic"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": lf["fused_frac_of_peak"],
            "traffic": _profiled_traffic(L, 32, shape),
            "traffic_source": f"rocprofv3 FETCH_SIZE + WRITE_SIZE per launch of this shape, "
                              f"{TRAFFIC_FILE.relative_to(ROOT)}",
            "kernel": ("k_fused_tile<float>" if shape["tiling"].startswith("LDS") else "k_fused<float>")
                      + " (E update + next link update per (site, dir)), FP32 storage",
            "launch_shape": shape,
            "bytes_per_launch": FUSED_BYTES[32] * V,
            "kernel_us": lf["fused_kernel_us"],
            "timed_launches": lf["fused_launches_timed"],
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    del g64, g32
    nrq.lattice_finalize()


if __name__ == "__main__":
    main()
