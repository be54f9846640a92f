"""Build the native parts of nrqcd_amd in-tree.

* ``nrqcd_amd/libnrqcd_amd.so`` -- the HIP library (gfx950) exporting the C ABI
  declared in ``include/nrqcd_amd.h``;
* ``nrqcd_amd/fortran/`` -- the ISO_C_BINDING Fortran modules (``lattice``,
  ``random``, ``gaugeconfiguration_su3``, ``nrqcd``, ``mpiinterface``) and the
  ``nrqcd_run`` driver, compiled with amdflang against the library.

Run ``python -m nrqcd_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
FSRC = PKG / "fortran"
LIB = PKG / "libnrqcd_amd.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))

HIP_SOURCES = ["nrq_core.hip", "nrq_gauge.hip", "nrq_ic.hip", "nrq_gaugefix.hip", "nrq_krylov.hip", "nrq_cn.hip", "nrq_cnn.hip", "nrq_spectrum.hip", "nrq_hamiltonian.hip", "nrq_equilibrium.hip"]
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # keep hipcc from emitting packed FP32 arithmetic
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(ROOT / "include"),
]
FORTRAN_MODULES = [
    "precision.f90",
    "nrq_capi.f90",
    "mpiinterface.f90",
    "lattice.f90",
    "random.f90",
    "gaugeconfiguration_su3.f90",
    "nrqcd.f90",
    "twfft.f90",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_hip(jobs: int = 8) -> Path:
    hipcc = str(ROCM / "bin" / "hipcc")
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    procs = []
    objs = []
    for src in HIP_SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        objs.append(obj)
        cmd = [hipcc, *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError("hipcc failed")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          "-L", str(ROCM / "lib"), "-lhipfft", "-Wl,-rpath," + str(ROCM / "lib")])
    os.replace(tmp, LIB)
    return LIB


def build_fortran() -> Path | None:
    """Compile the Fortran boundary modules and the driver with amdflang."""
    fc = shutil.which("amdflang") or str(ROCM / "bin" / "amdflang")
    if not Path(fc).exists():
        print("amdflang not found; skipping the Fortran boundary", file=sys.stderr)
        return None
    out = FSRC / "_build"
    out.mkdir(exist_ok=True)
    objs = []
    for src in FORTRAN_MODULES:
        obj = out / (Path(src).stem + ".o")
        _run([fc, "-O2", "-fPIC", "-module-dir", str(out), "-c", str(FSRC / src), "-o", str(obj)])
        objs.append(obj)
    exe = FSRC / "nrqcd_run"
    _run([fc, "-O2", "-module-dir", str(out), str(FSRC / "main.f90"), *map(str, objs),
          "-L", str(PKG), "-lnrqcd_amd", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + str(ROCM / "lib"),
          "-o", str(exe)])
    return exe


def main() -> None:
    build_hip()
    if (FSRC / "main.f90").exists():
        build_fortran()


if __name__ == "__main__":
    main()
