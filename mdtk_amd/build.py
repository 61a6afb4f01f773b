"""Builds the engine's shared library (libmdg.so) for gfx950 with hipcc.

The library is built in-tree so that it travels with the source tree; the
Python package loads it with ctypes.  FP64 kernels are compiled without FMA
contraction (-ffp-contract=off) so that every a*b+c rounds twice, as in the
reference's x86-64 build, and without packed-FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmdg.so")
SOURCES = ["mdg_api.hip", "mdg_kernels.hip", "mdg_force.hip", "mdg_brenner.hip", "mdg_tables.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-value", "-Wl,-z,defs",  # an unresolved symbol fails the link
         "-I" + os.path.join(HERE, "..", "include")]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    cmd = [HIPCC, *FLAGS, "-o", LIB + ".tmp", *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libmdg.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
