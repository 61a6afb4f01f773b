"""Builds the engine's shared library (libmdg.so) for gfx950 with hipcc.

The library is built in-tree so that it travels with the source tree; the
Python package loads it with ctypes.  FP64 kernels are compiled without FMA
contraction (-ffp-contract=off) so that every a*b+c rounds twice, as in the
reference's x86-64 build, and without packed-FP32 instructions.  FP32
division and square root use the hardware's approximate instructions (a few
ulp): FP32 arithmetic belongs to the mixed form alone, whose bar is a force
error, not bit parity.  The force kernels' file compiles twice: its FP64-form
kernels without contraction, its mixed-form kernels with FMA contraction
(-ffp-contract=fast).  The units compile in parallel into mdtk_amd/_build
and link into one library.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmdg.so")
# (source, object, extra flags)
UNITS = [("mdg_api.hip", "mdg_api", []), ("mdg_kernels.hip", "mdg_kernels", []),
         ("mdg_force.hip", "mdg_force64", ["-DMDG_FORCE_FP64"]),
         ("mdg_force.hip", "mdg_force32", ["-DMDG_FORCE_FP32", "-ffp-contract=fast"]),
         ("mdg_brenner.hip", "mdg_brenner", []), ("mdg_tables.cpp", "mdg_tables", [])]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-fno-hip-fp32-correctly-rounded-divide-sqrt",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-value",
         "-I" + os.path.join(HERE, "..", "include")]
LINK = ["--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,-z,defs"]  # an unresolved symbol fails the link


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
    return r.returncode


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    objs = [os.path.join(OBJ, o + ".o") for _, o, _ in UNITS]
    cmds = [[HIPCC, *FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
            for (src, _, extra), obj in zip(UNITS, objs)]
    with ThreadPoolExecutor(len(cmds)) as ex:
        rcs = list(ex.map(lambda c: _run(c, verbose), cmds))
    if any(rcs):
        raise RuntimeError("hipcc failed compiling the engine")
    if _run([HIPCC, *LINK, "-o", LIB + ".tmp", *objs], verbose):
        raise RuntimeError("hipcc failed linking libmdg.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
