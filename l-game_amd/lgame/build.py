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
SOURCES = ["lgame_capi.hip", "lgame_step.hip", "lgame_solve.hip"]
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


def build_all(verbose: bool = False) -> None:
    build_gpu(verbose)
    build_oracles(verbose)
