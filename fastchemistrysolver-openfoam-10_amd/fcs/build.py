"""Builds libfcs.so (the gfx950 HIP library behind include/fcs.h) and the
host harness of the C++ chemistry-model adapter (fcs_harness) in-tree."""
from __future__ import annotations

import os
import subprocess
import sys

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO = os.path.dirname(PKG_ROOT)
LIB = os.path.join(PKG_ROOT, "libfcs.so")
HARNESS = os.path.join(PKG_ROOT, "fcs_harness")
SRCS = [os.path.join(PKG_ROOT, "csrc", f) for f in ("fcs_lib.hip", "fcs_mlp.hip")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose: bool = False) -> str:
    """Compile csrc/*.hip into libfcs.so (always recompiles)."""
    cmd = [HIPCC, *HIP_FLAGS, "-I" + os.path.join(REPO, "include"),
           "-I" + os.path.join(PKG_ROOT, "csrc"), "-o", LIB + ".tmp", *SRCS]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "unknown target feature" not in l.lower())
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{out}")
    os.replace(LIB + ".tmp", LIB)
    build_harness(verbose)
    return LIB


def build_harness(verbose: bool = False) -> str:
    """The adapter harness: host C++ only (g++), linked against libfcs.so."""
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(REPO, "include"),
           os.path.join(PKG_ROOT, "csrc", "fcs_harness.cpp"), "-L" + PKG_ROOT, "-lfcs",
           "-Wl,-rpath,$ORIGIN", "-o", HARNESS + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"harness build failed ({r.returncode}):\n{r.stdout}")
    os.replace(HARNESS + ".tmp", HARNESS)
    return HARNESS


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
