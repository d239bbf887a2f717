"""Builds the HIP library (librocketcombustor.so) in-tree for gfx950.

    python -m rocketcombustor_amd.build

The kernels are compiled without floating-point contraction
(-ffp-contract=off) so that every FP64 expression rounds as the
reference's Fortran (x86-64 baseline, no FMA) does, and without packed FP32
instructions.  Symbols are hidden except the C-ABI of the header, and the
library's calls to functions it defines bind to its own definitions.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "librocketcombustor.so")
SOURCES = ["csrc/rc_api.hip", "csrc/rc_kernels.hip", "csrc/rc_backward_euler.hip", "csrc/rc_viscous.hip",
           "csrc/rc_mlp.hip", "csrc/rc_mesh.cpp"]
ARCH = os.environ.get("RC_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build_library(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-fvisibility=hidden", "-fvisibility-inlines-hidden", "-Wl,-Bsymbolic-functions",
           "-Wall", "-Wno-unused-function", "-I" + os.path.join(ROOT, "include"),
           "-o", tmp] + [os.path.join(PKG, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building " + LIB)
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = False) -> str:
    """The C oracle (test infrastructure) next to the package."""
    res = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("oracle build failed")
    return os.path.join(ROOT, "oracle", "librc_oracle.so")


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_oracle())
