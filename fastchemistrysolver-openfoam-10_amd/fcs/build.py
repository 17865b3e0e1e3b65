"""Builds libfcs.so (the gfx950 HIP library behind include/fcs.h) in-tree."""
from __future__ import annotations

import os
import subprocess
import sys

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO = os.path.dirname(PKG_ROOT)
LIB = os.path.join(PKG_ROOT, "libfcs.so")
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
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
