"""Builds nns_amd/libnbx.so (HIP for gfx950) with hipcc, in-tree.

    python -m nns_amd.build            # build (always recompiles)
    python -m nns_amd.build --runner   # also the C++ command-line runner nns_amd/nbx_runner
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnbx.so")
RUNNER = os.path.join(HERE, "nbx_runner")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NBX_ARCH", "gfx950")

# -ffp-contract=off: every fused multiply-add in the kernels is written out
# (fmaf), so FP results do not depend on compiler contraction.  Packed FP32
# instructions are disabled for all device code.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I", os.path.join(ROOT, "include"),
]

SOURCES = ["k_basic.hip", "k_radix.hip", "k_tree.hip", "k_force.hip", "k_adjoint.hip", "nbx_api.cpp", "nbx_batch.cpp",
           "nbx_ics.cpp", "k_probe.hip", "nbx_static.hip"]


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return out


def build(runner: bool = False, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    if runner:
        cmd = [HIPCC, *COMMON, "-x", "hip", os.path.join(CSRC, "nbx_runner.cpp"),
               os.path.join(CSRC, "barnes_hut.cpp"), "-o", RUNNER, f"-L{HERE}", "-lnbx",
               "-Wl,-rpath,$ORIGIN"]  # libnbx.so beside the runner, wherever the tree is
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"runner build failed:\n{r.stdout}\n{r.stderr}")
        if verbose:
            print(f"built {RUNNER}")
    return LIB


if __name__ == "__main__":
    build(runner="--runner" in sys.argv, verbose=True)
