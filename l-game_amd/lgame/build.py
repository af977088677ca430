"""Build of the in-tree native libraries.

* ``liblgame_gpu.so``: the HIP step for gfx950 (product).
* ``oracle/liboracle.so`` and, where the reference tree exists, ``oracle/_ref/libref.so``:
  the test oracles (built here so tests can run; never loaded by the product).
"""
from __future__ import annotations

import os
import subprocess

from .gpu import LIB_PATH, PKG_ROOT

REPO = os.path.dirname(PKG_ROOT)
CSRC = os.path.join(PKG_ROOT, "csrc")
SOURCES = ["lgame_capi.hip", "lgame_step.hip", "lgame_scalar.hip", "lgame_diff.hip", "lgame_solve.hip", "lgame_rad.hip",
           "lgame_mlp.hip", "lgame_grid.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # FP64 parity with the reference build: no FMA contraction
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build_gpu(verbose: bool = False) -> str:
    cmd = [HIPCC, *HIP_FLAGS, "-o", LIB_PATH + ".tmp", *[os.path.join(CSRC, s) for s in SOURCES]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed:\n{r.stderr}")
    os.replace(LIB_PATH + ".tmp", LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH}")
    return LIB_PATH


def build_oracles(verbose: bool = False) -> None:
    oracle_dir = os.path.join(REPO, "oracle")
    targets = ["oracle"]
    if os.path.isdir("/root/reference/src"):
        targets.append("ref")
    r = subprocess.run(["make", "-s", "-C", oracle_dir, *targets], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built oracle targets {targets}")


FORTRAN_DIR = os.path.join(PKG_ROOT, "fortran")
FORTRAN_BUILD = os.path.join(FORTRAN_DIR, "_build")
FORTRAN_DEMO = os.path.join(FORTRAN_BUILD, "lgame_driver_demo")


def build_fortran_demo(verbose: bool = False) -> str:
    """The iso_c_binding module and the demo Fortran time loop, linked against liblgame_gpu.so."""
    os.makedirs(FORTRAN_BUILD, exist_ok=True)
    fc = os.environ.get("FC", "flang")
    obj = os.path.join(FORTRAN_BUILD, "mo_lgame_gpu.o")
    cmds = [
        [fc, "-O2", "-c", os.path.join(FORTRAN_DIR, "mo_lgame_gpu.F90"), "-o", obj, "-module-dir", FORTRAN_BUILD],
        [fc, "-O2", "-I", FORTRAN_BUILD, os.path.join(FORTRAN_DIR, "lgame_driver_demo.F90"), obj,
         "-L", PKG_ROOT, "-llgame_gpu", "-Wl,-rpath,$ORIGIN/../..", "-o", FORTRAN_DEMO],
    ]
    for cmd in cmds:
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=FORTRAN_BUILD)
        if r.returncode != 0:
            raise RuntimeError(f"{' '.join(cmd)} failed:\n{r.stderr}")
    if verbose:
        print(f"built {FORTRAN_DEMO}")
    return FORTRAN_DEMO


def build_all(verbose: bool = False) -> None:
    build_gpu(verbose)
    build_oracles(verbose)
    build_fortran_demo(verbose)
