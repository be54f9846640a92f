"""Throughput of the real-time SU(3) leapfrog + NRQCD Crank-Nicolson hot path.

Workload (one GPU): 64^3 lattice, spacings (a0, a1, a2, a3) = (0.05, 1, 1, 1),
transverse-polarised box initial condition with Qs = 0.5, amplitude 1,
coupling 1, ranlxd seed 2; FP64 links and E; heavy quark and antiquark point
sources at lattice index 1, Ma = 2, c = (1, 0, 0, 0), every CN column solve
to an FP64 true relative residual of 1e-10.

A "step" is one leapfrog step of the whole lattice (3V link-updates).  The
timed region runs `--steps` leapfrog steps on the library's stream with the
inputs resident in HBM, and reads energy, Gauss law and plaquette every 10
steps as the workload logs them; CN steps are timed separately for each
solver: the plain FP64 GMRES(30) (PETSc's algorithm) and the fast ones.

Prints one JSON line.  Options: --gpus N (only 1), --steps K, --warmup W,
--size L, --cn-steps C, --no-cpu (skip the CPU baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "link-updates/s & CN steps/s at FP64 rel.res 1e-10: learned+mixed vs plain"
LOG_EVERY = 10  # steps between observables (energy, Gauss law, plaquette)
# Algorithmic HBM bytes per site of one mixed-Chebyshev CN step with 12 live
# fields when every column takes the planned 6 Chebyshev steps (two FP32
# corrections of 4 and 2 steps): right-hand side with y, x0 (FP64) and K psi
# (FP32) 2448; FP32 link copy 648; K pass 792; Chebyshev launches 792, 1080,
# 2232 (adds the correction to x0 in FP64) and 1656 (adds the second to psi);
# two FP64 true-residual passes 2 x 1872.
CN_MIXED_CHEB_BYTES = 2448.0 + 648 + 792 + 792 + 1080 + 2232 + 1656 + 2 * 1872
CN_MIXED_CHEB_ITS = 6.0
# Algorithmic HBM bytes per site of the leapfrog launches, FP64: the fused
# E + link pass reads and writes U and E of the three directions once; an
# Update of n steps is one links-only pass, n - 1 fused passes and one E-only
# pass (which leaves U and E at the same time for the observables).
FUSED_BYTES = 1248.0        # U 432 + E 192, read and written
LINKS_ONLY_BYTES = 1056.0   # U read and written, E read
E_ONLY_BYTES = 816.0        # U read, E read and written
HBM_PEAK_GBS = 8000.0
# HBM bytes per fused leapfrog launch measured with rocprofv3 (FETCH_SIZE and
# WRITE_SIZE, each calibrated on the copy kernel), kept with the profile
TRAFFIC_FILE = ROOT / "profiles" / "leapfrog_fused_64cubed_fp64_traffic.txt"


TILINGS = ("consecutive sites", "z-sweep tiles of x-rows", "z-sweep 16-wide tiles")


def _shape_key(shape: dict) -> str:
    return (f"bs={shape['threads_per_block']} xcd={shape['xcd_remap']} tiling={TILINGS.index(shape['tiling'])} "
            f"streaming={int(shape['streaming'])} two_row={int(shape['two_row_neighbours'])}")


def _measured_traffic(size: int, shape: dict):
    """HBM bytes per fused launch of this launch shape from the committed
    rocprof summary, or None if that shape was not measured."""
    if size != 64 or not TRAFFIC_FILE.exists():
        return None
    key = f"traffic_bytes_per_launch[{_shape_key(shape)}]="
    for line in TRAFFIC_FILE.read_text().splitlines():
        if line.startswith(key):
            return float(line[len(key):])
    return None


def _cpu_baseline(size: int, seconds: float) -> dict:
    """Oracle leapfrog (OpenMP FP64 C restatement) on a bounded sample of the
    same workload, energy, Gauss law and plaquette every LOG_EVERY steps: the
    number of 10-step chunks is chosen to fill about `seconds`."""
    from oracle.pyoracle import Oracle

    orc = Oracle()
    orc.lattice_init((size, size, size), (0.05, 1.0, 1.0, 1.0))
    U, E = orc.box_init(2, 0.5, 1.0, 1.0)

    def chunk():
        orc.leapfrog(U, E, 1.0, LOG_EVERY)
        return (orc.energy(U, E), orc.gauss_ref(U, E), orc.gauss_max(U, E), orc.mean_plaquette(U))

    t0 = time.perf_counter()
    chunk()  # warm-up and calibration
    one = time.perf_counter() - t0
    chunks = max(1, int(seconds / max(one, 1e-6)))
    t0 = time.perf_counter()
    for _ in range(chunks):
        chunk()
    dt = time.perf_counter() - t0
    steps = chunks * LOG_EVERY
    return {
        "value": 3 * size ** 3 * steps / dt,
        "unit": "link-updates/s",
        "cores": orc.num_threads(),
        "kind": "port",
        "sample": f"oracle leapfrog FP64 with energy, Gauss law and plaquette every {LOG_EVERY} steps, {size}^3, "
                  f"{steps} steps after one warm-up chunk ({dt:.1f} s)",
    }


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=64)
    ap.add_argument("--cn-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for {args.gpus}); the lattice is not "
              "sharded across GPUs", file=sys.stderr)
        sys.exit(2)

    import nrqcd_amd as nrq

    L = args.size
    V = L ** 3
    nrq.lattice_init((L, L, L), (0.05, 1.0, 1.0, 1.0))
    nrq.random_init(2)
    g = nrq.GaugeConfiguration(nrq.NRQ_PRECISION_FP64)
    g.TransversePolarisedOccupiedInit_Box(0.5, 1.0, 1.0)
    obs0 = g.observables()

    # ---- leapfrog ------------------------------------------------------
    # The workload logs energy, Gauss law and plaquette every 10 steps: the
    # timed region runs the steps in chunks of (at most) 10, each followed by
    # a logged observables computation that stays on the GPU (read once at
    # the end).  HIP events around each chunk's Update give the leapfrog
    # kernel time inside the same region, for the roofline; a pair of event
    # slots is read back PAIRS chunks after it was recorded, so the host
    # never drains the queue.
    PAIRS = 7  # event slots 2..15
    g.Update(1.0, args.warmup)
    nrq.synchronize()
    nrq.event_record(0)
    t0 = time.perf_counter()
    lf_ms, lf_bytes, done, chunk = 0.0, 0.0, 0, 0
    while done < args.steps:
        n = min(LOG_EVERY, args.steps - done)
        k = chunk % PAIRS
        if chunk >= PAIRS:
            lf_ms += nrq.event_elapsed_ms(2 + 2 * k, 3 + 2 * k)
        nrq.event_record(2 + 2 * k)
        g.Update(1.0, n)
        nrq.event_record(3 + 2 * k)
        g.log_observables()
        lf_bytes += V * (FUSED_BYTES * (n - 1) + LINKS_ONLY_BYTES + E_ONLY_BYTES)
        done += n
        chunk += 1
    nrq.event_record(1)
    ms = nrq.event_elapsed_ms(0, 1)
    wall = time.perf_counter() - t0
    for c in range(max(0, chunk - PAIRS), chunk):
        k = c % PAIRS
        lf_ms += nrq.event_elapsed_ms(2 + 2 * k, 3 + 2 * k)
    log = g.read_log()
    obs1 = dict(zip(("energy", "gauss_ref", "gauss_max", "plaquette"), log[-1]))
    shape = g.launch_shape()
    lu_per_s = 3 * V * args.steps / (ms * 1e-3)
    ms_step = ms / args.steps
    # the leapfrog launches' algorithmic bytes over their event time
    achieved = lf_bytes / (lf_ms * 1e-3) / 1e9

    # ---- CN ----------------------------------------------------------------
    nrq.nrqcd_init(2.0, [1, 0, 0, 0], 1.0, 1e-10)
    cn = {}
    for name, kind in (("plain_fp64_gmres", nrq.NRQ_SOLVER_GMRES_FP64),
                       ("mixed_gmres", nrq.NRQ_SOLVER_GMRES_MIXED),
                       ("fp64_chebyshev", nrq.NRQ_SOLVER_CHEBYSHEV_FP64),
                       ("mixed_chebyshev", nrq.NRQ_SOLVER_CHEBYSHEV_MIXED)):
        nrq.nrqcd_set_solver(kind, 30, 10000)
        h = nrq.NRQCDField()
        h.InitSinglePoint(1, 1)
        h.Update(g, 2.0, [1, 0, 0, 0])  # warm-up (workspace allocation)
        nrq.synchronize()
        its = 0
        solves = 0
        worst = 0.0
        nrq.event_record(2)
        for _ in range(args.cn_steps):
            st = h.Update(g, 2.0, [1, 0, 0, 0])
            its += st.iterations
            solves += st.solves
            worst = max(worst, st.max_true_relres)
        nrq.event_record(3)
        cn_ms = nrq.event_elapsed_ms(2, 3)
        nq, na = h.norms()
        cn[name] = {"cn_steps_per_s": args.cn_steps / (cn_ms * 1e-3), "ms_per_step": cn_ms / args.cn_steps,
                    "column_solves_per_s": solves / (cn_ms * 1e-3),
                    "iterations_per_column_solve": its / max(solves, 1), "worst_true_relres": worst,
                    "norms": [nq, na]}
        del h
    nrq.nrqcd_finalize()
    copy_gbs = nrq.bandwidth_probe(1 << 30, 20)

    cpu = None
    # HBM roofline of the fastest CN step, from its fixed launch sequence
    mc = cn["mixed_chebyshev"]
    cn_roof = None
    if mc["iterations_per_column_solve"] == CN_MIXED_CHEB_ITS:
        gbs = CN_MIXED_CHEB_BYTES * V / (mc["ms_per_step"] * 1e-3) / 1e9
        cn_roof = {"solver": "mixed_chebyshev", "bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": gbs / HBM_PEAK_GBS, "bytes_per_step": CN_MIXED_CHEB_BYTES * V}
    if not args.no_cpu:
        cpu = _cpu_baseline(L, args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": lu_per_s,
        "unit": "link-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{L}^3 SU(3) leapfrog (FP64) + NRQCD CN point source, box IC Qs=0.5 seed 2",
            "lattice": [L, L, L],
            "spacings": [0.05, 1.0, 1.0, 1.0],
            "cn_solvers": "GMRES(30) FP64 (plain), FP32-GMRES/FP64-IR, Chebyshev FP64, FP32-Chebyshev/FP64-IR",
            "mass": 2.0,
        },
        "cn_steps_per_s": cn["mixed_chebyshev"]["cn_steps_per_s"],
        "cn_steps_per_s_plain": cn["plain_fp64_gmres"]["cn_steps_per_s"],
        "cn": cn,
        "cn_speedup_vs_plain": {k: v["cn_steps_per_s"] / cn["plain_fp64_gmres"]["cn_steps_per_s"] for k, v in cn.items()},
        "cn_roofline": cn_roof,
        "logged_rows": int(log.shape[0]),
        "energy_drift_rel": (obs1["energy"] - obs0["energy"]) / obs0["energy"],
        "gauss_max_drift": obs1["gauss_max"] - obs0["gauss_max"],
        "wall_s_leapfrog": wall,
        "hbm_copy_probe_gbs": copy_gbs,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": _measured_traffic(L, shape),
            "kernel": "k_fused<double> (E update + next link update per (site, dir))",
            "launch_shape": shape,
            "bytes_per_launch": FUSED_BYTES * V,
            "achieved_over": "all leapfrog launches of the timed region (fused, links-only, E-only)",
            "leapfrog_ms_per_step": lf_ms / args.steps,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    nrq.lattice_finalize()


if __name__ == "__main__":
    main()
