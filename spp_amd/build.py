"""Build the engine library in-tree: hipcc for gfx950, one shared object, C ABI only."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libspp_engine.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: products and sums round like the reference's x86-64 build (no FMA), so the
# Verlet predicate sees bit-identical distances.  Packed FP32 is disabled for all HIP code.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build_engine(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, "engine.hip"), os.path.join(CSRC, "surrogate.hip")]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("hipcc failed for libspp_engine.so")
    os.replace(tmp, LIB)
    return LIB


DEMOS = ("box_demo", "vesicle_demo", "state_demo")


def build_demo(verbose: bool = False) -> list:
    """The C++ Box facade drivers (host code only, linked against the engine library)."""
    outs = []
    for name in DEMOS:
        out = os.path.join(HERE, name)
        cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, name + ".cpp"),
               "-L", HERE, "-lspp_engine", "-Wl,-rpath,$ORIGIN", "-o", out]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stderr)
            raise RuntimeError("g++ failed for " + name)
        outs.append(out)
    return outs


if __name__ == "__main__":
    print(build_engine(verbose=True))
    print(build_demo(verbose=True))
