"""Builds libm1lya.so (the HIP/gfx950 shared library) in-tree.

    python -m lydion_amd.build          # or lydion_amd.build.build()

hipcc cross-compiles for gfx950 without a GPU.  FP contraction is disabled so
that the FP64 step rounds exactly as its formulas are written, and packed FP32
instructions are disabled (-packed-fp32-ops).  Machine-level loop-invariant code
motion is disabled: in the long cell-marching loops it hoists every constant of
the inlined exp / division sequences into a VGPR held across the loop (about 30
VGPRs in the batched step), which costs occupancy or spills; rematerialising the
constants inside the loop is scalar work that is cheaper.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libm1lya.so")
SOURCES = ["m1lya_kernels.hip", "m1lya_api.hip", "m1lya_batch.hip", "m1lya_adjoint.hip", "m1lya_noalloc.hip"]
ARCH = os.environ.get("M1LYA_OFFLOAD_ARCH", "gfx950")

COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-disable-machine-licm",
    "-Wall",
    "-Wno-unused-function",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libm1lya")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    lines = [ln for ln in proc.stdout.splitlines() if "is not a recognized feature for this target" not in ln]
    if proc.returncode != 0:
        raise RuntimeError("command failed:\n  " + " ".join(cmd) + "\n" + "\n".join(lines))
    if lines:
        sys.stderr.write("\n".join(lines) + "\n")


def build(verbose: bool = False) -> str:
    """Compiles every HIP source for gfx950 and links libm1lya.so; returns its path."""
    hipcc = _hipcc()
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([hipcc, *COMMON_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for cmd, fut in zip(cmds, [ex.submit(_run, c) for c in cmds]):
            if verbose:
                print(" ".join(cmd))
            fut.result()
    tmp = LIB + ".tmp"
    link = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(link))
    _run(link)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
