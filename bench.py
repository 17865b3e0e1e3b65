#!/usr/bin/env python3
"""Benchmark of the gfx950 per-cell chemistry integrator.

One step = FastChemistryModel::solve(deltaT) over the whole batch of cells
with inputs resident in HBM: plain OptRodas34, FP64, the 53-species /
325-reaction synthetic mechanism, 65,536 cells on a 1-D mixture-fraction
profile, deltaT = 1e-6 s, absTol 1e-8, relTol 1e-4, initial chemical
time step 1e-7 s.

Prints ONE JSON line (rank 0).  The CPU baseline is the oracle (a scalar C
restatement of the reference algorithm) timed on a bounded, seeded sample
of the same cells on one host core.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fastchemistrysolver-openfoam-10_amd"))

METRIC = "cells·dt/s at 53-species mech (learned-seeded vs plain GPU Rodas34, ≤1e-4 err)"
FP64_PEAK_TFLOPS = 78.6      # MI355X dense FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0


def cpu_baseline(mech, field, dt, seconds=15.0):
    """Times the oracle on one core over a seeded random sample of cells,
    growing the sample until ~`seconds` of CPU work."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    o = Oracle(mech)
    rng = np.random.default_rng(7)
    order = rng.permutation(field.n_cells)
    n, used, cells = 8, 0.0, 0
    while used < seconds and cells < field.n_cells:
        idx = order[cells:cells + n]
        t0 = time.perf_counter()
        o.solve(field.subset(idx), dt, threads=1)
        used += time.perf_counter() - t0
        cells += len(idx)
        n = min(2 * n, 256)
    return dict(value=cells / used, unit="cells·dt/s", cores=1, kind="port",
                sample=f"{cells} of {field.n_cells} cells drawn uniformly at random (seed 7), "
                       f"{used:.1f} s on one core, plain OptRodas34 oracle (C, -O3)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=65536)
    ap.add_argument("--solver", default="OptRodas34")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); cells are independent, "
              "so the path shards by cell ranges with no collective on the data path.", file=sys.stderr)
        sys.exit(2)

    from fcs import lib
    from fcs.flops import attempt_flops, hbm_bytes_per_cell
    from fcs.mechanism import config_field, config_mechanism

    mech = config_mechanism(1)
    field = config_field(1, mech, n_cells=args.cells)
    dt = 1e-6
    gm = lib.Mech(mech)
    solver = lib.Solver(gm, solver=args.solver, jacobian="exact", absTol=1e-8, relTol=1e-4)
    batch = lib.Batch(solver, field)

    for _ in range(args.warmup):
        batch.run(dt, timed=True)
    batch.profile(True)     # HIP event pair around every step launch, on the solver's stream
    launches, step_ms = 0, 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run(dt, timed=True)
        n, ms = batch.stats()
        launches += n
        step_ms += ms
    wall = time.perf_counter() - t0
    out = batch.download()
    nAtt = int(out["nAttempts"].sum())          # every step re-integrates the same inputs
    bad = int(((out["status"] & 6) != 0).sum())

    ms_step = 1e3 * wall / args.steps
    value = args.cells / (wall / args.steps)
    fpa = attempt_flops(mech, args.solver)
    avg_launch_s = 1e-3 * step_ms / max(launches, 1)
    flops_launch = fpa * nAtt * args.steps / max(launches, 1)
    achieved = flops_launch / avg_launch_s / 1e12
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cells·dt/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain {args.solver}, 53-species/325-reaction synthetic mechanism, "
                               f"{args.cells} cells (1-D mixture-fraction profile), deltaT 1e-6 s",
                   "cells": args.cells, "species": mech.n_species, "reactions": mech.n_reactions,
                   "solver": args.solver, "jacobian": "exact", "absTol": 1e-8, "relTol": 1e-4,
                   "attempts_per_step": nAtt, "failed_cells": bad},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "k_step<56,Rodas34> (FP64 VALU; one Rosenbrock attempt per active cell)",
                     "flops_per_attempt": fpa,
                     "launches": launches,
                     "avg_launch_ms": 1e3 * avg_launch_s,
                     "step_kernel_ms_per_step": step_ms / args.steps,
                     "hbm_GBs_algorithmic": hbm_bytes_per_cell(mech) * args.cells * args.steps / (1e-3 * step_ms) / 1e9},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(mech, field, dt)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
