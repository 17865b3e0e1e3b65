"""Builds libfcs.so (the gfx950 HIP library behind include/fcs.h) and the
host harness of the C++ chemistry-model adapter (fcs_harness) in-tree."""
from __future__ import annotations

import os
import re
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO = os.path.dirname(PKG_ROOT)
LIB = os.path.join(PKG_ROOT, "libfcs.so")
# profiling variant: the same library with the per-phase cycle counters of
# the attempt compiled in (-DFCS_PHASE_TIMING, fcs_phase_read exported)
LIB_PHASES = os.path.join(PKG_ROOT, "libfcs_phases.so")
HARNESS = os.path.join(PKG_ROOT, "fcs_harness")
CSRC = os.path.join(PKG_ROOT, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def padded_sizes() -> list:
    """The padded system sizes fcs_kernels.h lists (FCS_FOR_EACH_NP and
    FCS_FOR_EACH_NP_LARGE): one object of fcs_kernels_np.hip each."""
    with open(os.path.join(CSRC, "fcs_kernels.h")) as f:
        text = f.read()
    sizes = []
    for name in ("FCS_FOR_EACH_NP", "FCS_FOR_EACH_NP_LARGE"):
        m = re.search(r"#define " + name + r"\(X\)((?: X\(\d+\))+)", text)
        if not m:
            raise RuntimeError(f"{name} not found in fcs_kernels.h")
        sizes += [int(v) for v in re.findall(r"X\((\d+)\)", m.group(1))]
    return sizes


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "unknown target feature" not in l.lower())
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{' '.join(cmd)}\n{out}")


def build(verbose: bool = False, jobs: int | None = None, phase_timing: bool = False) -> str:
    """Compile csrc/*.hip into libfcs.so (always recompiles): the host side,
    the network kernels and one object per padded size, compiled in
    parallel, then linked.  phase_timing builds the profiling variant
    libfcs_phases.so instead."""
    inc = ["-I" + os.path.join(REPO, "include"), "-I" + CSRC]
    if phase_timing:
        inc.append("-DFCS_PHASE_TIMING")
    out = LIB_PHASES if phase_timing else LIB
    with tempfile.TemporaryDirectory(prefix="fcs_build_") as tmp:
        units = [(os.path.join(CSRC, "fcs_lib.hip"), [], os.path.join(tmp, "fcs_lib.o")),
                 (os.path.join(CSRC, "fcs_mlp.hip"), [], os.path.join(tmp, "fcs_mlp.o"))]
        units += [(os.path.join(CSRC, "fcs_kernels_np.hip"), [f"-DFCS_NP={n}"], os.path.join(tmp, f"fcs_np{n}.o"))
                  for n in padded_sizes()]
        cmds = [[HIPCC, *HIP_FLAGS, *inc, *defs, "-c", src, "-o", obj] for src, defs, obj in units]
        jobs = jobs or max(1, min(len(cmds), os.cpu_count() or 1, 8))
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            for f in [ex.submit(_run, c, verbose) for c in cmds]:
                f.result()
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-o", out + ".tmp", *[obj for _, _, obj in units]], verbose)
    os.replace(out + ".tmp", out)
    if not phase_timing:
        build_harness(verbose)
    return out


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
    print(build(verbose=True, phase_timing="--phase-timing" in sys.argv[1:]))
    sys.exit(0)
