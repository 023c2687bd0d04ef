"""Builds the in-tree HIP library libsa_hip.so for gfx950 with hipcc.

    python -m simulated_annealing_amd.build        # or __graft_entry__.build()

The library holds the sweep/measure/swap kernels, the C ABI of include/sa_hip.h and
the C++ drivers behind it.  It is rebuilt every time this runs.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libsa_hip.so")
CLI = os.path.join(PKG, "sa_main")  # command-line driver with the reference's flags

ARCH = os.environ.get("SA_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # keep the compiler from emitting packed FP32 instructions (none are wanted here)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]

SOURCES = ["sa_ladder.hip", "sa_batch.hip", "sa_noalloc.hip", "sa_serve.cpp", "sa_drivers.cpp"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = True) -> str:
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIP_FLAGS, "-shared", "-I", INCLUDE, "-I", CSRC,
           *[os.path.join(CSRC, s) for s in SOURCES], "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    # the command-line driver: host C++ over the library
    cli = [hipcc(), "-O2", "-std=c++17", "-Wall", "-Wno-unknown-pragmas", "-I", INCLUDE,
           os.path.join(PKG, "cli", "sa_main.cpp"), "-L", PKG, "-lsa_hip", "-Wl,-rpath,$ORIGIN", "-o", CLI + ".tmp"]
    if verbose:
        print(" ".join(cli), flush=True)
    subprocess.check_call(cli)
    os.replace(CLI + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
