"""Build the MI355X LES library (liblesgpu.so) and the Fortran host driver.

hipcc compiles every .hip source for gfx950 (cross-compilation works without
a GPU); flang builds the iso_c_binding host modules and driver program.
Outputs stay in-tree so they travel with the source tree.
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG / "csrc"
FSRC = PKG / "fortran"
LIB = PKG / "liblesgpu.so"
FDRIVER = PKG / "les_driver"

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # every multiply and add rounded separately: results reproducible and
    # comparable operation-for-operation with the FP64 CPU restatement
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result",
]

SOURCES = ["les_gpu.hip"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"command failed: {' '.join(map(str, cmd))}")
    return r


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and pathlib.Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def flang() -> str | None:
    for c in ("/opt/rocm/lib/llvm/bin/flang", shutil.which("flang"), shutil.which("amdflang")):
        if c and pathlib.Path(c).exists():
            return c
    return None


def build_lib() -> pathlib.Path:
    objs = []
    bdir = PKG / "_build"
    bdir.mkdir(exist_ok=True)
    for s in SOURCES:
        src = CSRC / s
        if not src.exists():
            continue
        obj = bdir / (src.stem + ".o")
        _run([hipcc(), *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs])
    tmp.replace(LIB)
    return LIB


def build_fortran() -> pathlib.Path | None:
    fc = flang()
    if fc is None or not FSRC.exists():
        return None
    bdir = PKG / "_build" / "fortran"
    bdir.mkdir(parents=True, exist_ok=True)
    order = ["les_gpu_abi.f90", "parameters_IO.f90", "system_variables.f90", "compute_modules.f90",
             "running_diagnostics.f90", "IO_statistics.f90", "tke_budget1D.f90", "IO_main.f90",
             "initial_conditions.f90", "LES_engine.f90",
             "main.f90"]
    srcs = [str(FSRC / f) for f in order if (FSRC / f).exists()]
    if not srcs:
        return None
    _run([fc, "-O2", "-J", str(bdir), *srcs, "-L", str(PKG), "-llesgpu",
          f"-Wl,-rpath,{PKG}", "-o", str(FDRIVER)], cwd=str(bdir))
    return FDRIVER


def build_all() -> None:
    build_lib()
    build_fortran()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
