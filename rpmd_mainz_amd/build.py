"""Build librpmd_mi355x.so (gfx950) in-tree with hipcc.

Every HIP translation unit under csrc/ is compiled for --offload-arch=gfx950 with packed FP32
instructions disabled, then linked into rpmd_mainz_amd/librpmd_mi355x.so next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "librpmd_mi355x.so")
BUILD = os.path.join(HERE, "_build")
SOURCES = ["k_forces.hip", "k_step.hip", "k_rigid.hip", "capi.hip", "ipi.cpp"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-munsafe-fp-atomics", "-I" + os.path.join(HERE, "..", "include")]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}", file=sys.stderr)
    return OUT


FORTRAN_DIR = os.path.join(HERE, "fortran")
HARNESS = os.path.join(FORTRAN_DIR, "rpmd_forces_harness")


def build_fortran(verbose: bool = True) -> str | None:
    """Compile the ISO_C_BINDING layer (fortran/rpmd_hip.f90) and the Fortran host harness of the
    tests (tests/fortran/harness.f90) against librpmd_mi355x.so with amdflang, when present."""
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    if fc is None:
        return None
    os.makedirs(BUILD, exist_ok=True)
    obj = os.path.join(BUILD, "rpmd_hip.o")
    r = subprocess.run([fc, "-O2", "-c", os.path.join(FORTRAN_DIR, "rpmd_hip.f90"), "-J", BUILD, "-o", obj],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed for rpmd_hip.f90:\n{r.stderr}")
    harness_src = os.path.join(HERE, "..", "tests", "fortran", "harness.f90")
    r = subprocess.run([fc, "-O2", harness_src, obj, "-J", BUILD, "-L", HERE, "-lrpmd_mi355x",
                        "-Wl,-rpath,$ORIGIN/..", "-o", HARNESS], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang link of the harness failed:\n{r.stderr}")
    if verbose:
        print(f"built {HARNESS}", file=sys.stderr)
    return HARNESS


if __name__ == "__main__":
    build()
    build_fortran()
