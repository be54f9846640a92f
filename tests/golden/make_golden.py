"""Write tests/golden/ref_vectors.npz from the reference's own compiled modules.

The reference's pure Fortran modules (ranlxd_generator, su3, su2,
matrixoperations) are compiled from /root/reference/src by
oracle/build_ref.sh into oracle/_ref/libnrqcd_ref.so.  This script calls them
and stores their outputs as small fixtures, so the oracle and the product can
be checked against the reference's arithmetic where /root/reference is absent.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.pyoracle import RefModules  # noqa: E402

OUT = Path(__file__).resolve().parent / "ref_vectors.npz"


def main() -> None:
    ref = RefModules()
    data = {}
    # ranlxd streams (ranlxd_generator.f90): levels 1 and 2, the seeds of the
    # configured workloads, 3000 numbers each and the 25-word state after them
    for level in (1, 2):
        for seed in (1, 2, 3, 100, 131, 2147483647):
            ref.ranlxd_init(level, seed)
            data[f"ranlxd_l{level}_s{seed}"] = ref.ranlxd(3000)
            data[f"ranlxd_state_l{level}_s{seed}"] = ref.ranlxd_get()
    # generator tables (su3.f90:97-131, su2.f90:68-88)
    data["su3_generators"] = ref.su3_generators()
    data["su2_generators"] = ref.su2_generators()
    # group exponential GetGroupExp(alpha) over several norm scales
    rng = np.random.default_rng(20240601)
    alphas = []
    for scale in (1e-6, 1e-3, 0.05, 0.3, 1.0, 3.0, 10.0):
        for _ in range(8):
            alphas.append(rng.normal(size=8) * scale)
    alphas = np.array(alphas)
    data["exp_alpha"] = alphas
    data["exp_U"] = np.array([ref.group_exp(a) for a in alphas])
    # algebra maps
    mats = rng.normal(size=(16, 3, 3)) + 1j * rng.normal(size=(16, 3, 3))
    data["alg_M"] = mats
    data["alg_coord"] = np.array([[ref.algebra_coordinate(a, m) for a in range(1, 9)] for m in mats])
    data["alg_trace"] = np.array([[ref.trace_with_generator(a, m) for a in range(1, 9)] for m in mats])
    data["alg_matrix"] = np.array([ref.algebra_matrix(a) for a in alphas[:16]])
    # Kronecker product (matrixoperations.f90:81-104), colour (x) spin
    A = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    B = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
    data["kron_A"], data["kron_B"], data["kron_AB"] = A, B, ref.kron(A, B)
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
