"""Builds libbtk.so (HIP kernels + C ABI) in-tree with hipcc for gfx950.

    python -m barytreek_amd.build        # or barytreek_amd.build.build()

Objects go to barytreek_amd/_build/, the library to barytreek_amd/libbtk.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbtk.so")
DRIVER = os.path.join(HERE, "btk_driver")
REPO = os.path.dirname(HERE)

SOURCES = [
    "btk_runtime.hip",
    "btk_tree.hip",
    "btk_traversal.hip",
    "btk_passes.hip",
    "btk_interact.hip",
    "btk_ik_2d_pot.hip",
    "btk_ik_2d_grad.hip",
    "btk_ik_3d_pot.hip",
    "btk_ik_3d_vel.hip",
    "btk_ik_3d_grad.hip",
    "btk_mrhs.hip",
    "btk_mutual.hip",
    "btk_mac.hip",
    "btk_api.hip",
]

ARCH = "gfx950"

# -ffp-contract=off: geometry, node and basis arithmetic round exactly as
# written (bit-exact trees and lists); hot loops spell out fma().
# -packed-fp32-ops off: no packed FP32 instructions are generated.
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++20", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I", os.path.join(REPO, "include"),
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # namelist-driven driver over the C++ drop-in layer (cpp/barytreek.hpp)
    cmd = [hipcc(), "-O2", "-std=c++20", "-I", os.path.join(REPO, "include"), "-I", os.path.join(HERE, "cpp"),
           os.path.join(HERE, "cpp", "btk_driver.cpp"), "-o", DRIVER + ".tmp", "-L", HERE, "-lbtk",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(DRIVER + ".tmp", DRIVER)
    if verbose:
        print(f"built {LIB} and {DRIVER}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
