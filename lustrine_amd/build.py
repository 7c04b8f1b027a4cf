"""Builds lustrine_amd/liblustrine_gpu.so for gfx950 with hipcc, in-tree.

Flags that fix the arithmetic:
  -ffp-contract=off                           no a*b+c fusion: one rounding per op
  -fhip-fp32-correctly-rounded-divide-sqrt    IEEE-754 fp32 '/' and sqrtf
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                              no packed-fp32 instruction selection
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblustrine_gpu.so")
SOURCES = ["lus_kernels.hip", "lus_pbf.hip", "lus_mlp.hip", "lus_noalloc.hip", "lus_fluid.hip", "lus_capi.cpp"]
ARCH = os.environ.get("LUS_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def command(out: str = LIB) -> list:
    return [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
        *[os.path.join(CSRC, s) for s in SOURCES],
        "-o", out,
    ]


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building liblustrine_gpu.so")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
