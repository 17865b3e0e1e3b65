#!/usr/bin/env python3
"""Benchmark of the gfx950 per-cell chemistry integrator.

One step = FastChemistryModel::solve(deltaT) over the whole batch of cells
with inputs resident in HBM (fcs_batch_run), FP64 Rodas34, the
53-species / 325-reaction synthetic mechanism, deltaT = 1e-6 s,
absTol 1e-8, relTol 1e-4, initial chemical time step 1e-7 s.

Default workload (the metric's): learned-seeded Rodas34 on the 512 x 512
counterflow field (262,144 cells: 45 % cold cells below Treact = 500 K,
40 % near-equilibrium products, 15 % flame front), step launches replayed
from a hipGraph.  The network is trained first (not timed) by the
library's own BF16-MFMA kernels on 2 M samples the plain GPU integrator
produces from other seeded 256 x 256 counterflow fields (56 of them).  The same field is also
run with the plain integrator (no network), and the 65,536-cell
mixture-fraction field of the plain-integrator configuration is run too;
their rates are reported beside `value`.  With the CPU baseline, the error
of the plain, learned and seed-mode runs on a stratified sample of product
and flame-front cells is measured against the oracle's relTol 1e-10 /
absTol 1e-14 solve of those cells (`error_vs_oracle_tight`).

Prints ONE JSON line.  The CPU baseline is the oracle (a scalar C
restatement of the reference algorithm) timed on a bounded, stratified
sample of the same field on one host core, and the same sample again on
all host cores this process may use (at most 16) as `all_cores`.
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
DT = 1e-6
TREACT = 500.0
GRAPH_LAUNCHES = 24


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    """Host CPU model name (/proc/cpuinfo, as lscpu prints it) and the core
    counts: nproc of the machine and the cores this process may use."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    name = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(model=name, nproc=os.cpu_count(), usable=len(os.sched_getaffinity(0)))


def cpu_baseline(mech, field, n_cells=4096, gpu=None, runs=3):
    """Times the oracle's AVX2 build (-O3 -mavx2 -mfma, libmvec exp/log) on a
    stratified random sample of n_cells cells of the field, drawn in the
    field's class proportions so that cells/s of the sample is the field's
    cells·dt/s: once on one core, then the median of `runs` runs on every
    usable host core (at most 16).  With gpu (a download() dict of the GPU
    run over the whole field) it also reports the largest per-species |dY|
    and |dT| between the CPU and GPU results on the sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    o = Oracle(mech, variant="avx2")
    rng = np.random.default_rng(7)
    classes, counts = np.unique(field.cls, return_counts=True)
    n_cells = min(n_cells, field.n_cells)
    take = {int(c): int(round(n_cells * k / field.n_cells)) for c, k in zip(classes, counts)}
    idx = np.sort(np.concatenate([rng.choice(np.nonzero(field.cls == c)[0], take[c], replace=False)
                                  for c in take]))
    sub = field.subset(idx)
    warm = field.subset(np.concatenate([np.nonzero(field.cls == c)[0][:2] for c in take]))
    o.solve(warm, DT, threads=1, Treact=TREACT)
    t0 = time.perf_counter()
    r1 = o.solve(sub, DT, threads=1, Treact=TREACT)
    t1 = time.perf_counter() - t0
    threads = host_threads()
    walls = []
    for _ in range(runs):
        t0 = time.perf_counter()
        o.solve(sub, DT, threads=threads, Treact=TREACT)
        walls.append(time.perf_counter() - t0)
    tm = float(np.median(walls))
    desc = ", ".join(f"class {c}: {take[c]}" for c in sorted(take))
    out = dict(value=idx.size / t1, unit="cells·dt/s", cores=1, kind="port",
               sample=f"stratified random sample of {idx.size} cells of the {field.n_cells}-cell field in its "
                      f"class proportions ({desc}; 0 cold, 1 product, 2 flame front), {t1:.1f} s on one core; "
                      "plain OptRodas34 oracle built -O3 -mavx2 -mfma -ffast-math with libmvec 4-wide exp/log "
                      "(oracle/liboracle_avx2.so), after a warm-up solve of 2 cells per class",
               host=cpu_model(),
               all_cores=dict(value=idx.size / tm, cores=threads,
                              sample=f"the same sample, median of {runs} runs ({', '.join(f'{w:.2f}' for w in walls)} "
                                     f"s wall) on {threads} threads (OpenMP over cells, dynamic schedule)"))
    if gpu is not None:
        rho = r1["rho0"]
        dY = np.abs(r1["RR"] - gpu["RR"][:, idx]) * DT / rho[None, :]
        Yc = np.abs(sub.Y + r1["RR"] * DT / rho[None, :])
        ok = (r1["status"] & 6) == 0
        over = (dY > np.maximum(1e-6, 1e-4 * Yc))[:, ok]
        dT = float(np.abs(r1["Tend"] - gpu["Tend"][idx])[ok].max())
        out["vs_gpu"] = dict(max_abs_dY=float(dY[:, ok].max()), entries_over_dY_bar=int(over.sum()),
                             max_abs_dT_K=dT, holds=bool(not over.any() and dT <= 1e-3),
                             bar="per-species |dY| <= 1e-6 absolute or <= 1e-4 relative, |dT| <= 1e-3 K")
    return out


def counted_flops(mech, solver="OptRodas34", jacobian="exact"):
    """FP64 operations of one attempt: the frozen count of the op-counting
    oracle build (tests/golden/flops.json, rank-1 dC/dY conversion) for the
    benchmark mechanisms, else the hand model of fcs/flops.py.  Returns
    (flops, source)."""
    from fcs.flops import attempt_flops
    name = f"synthetic_{mech.n_species}sp_{mech.n_reactions}rxn"
    try:
        with open(os.path.join(ROOT, "tests", "golden", "flops.json")) as f:
            return json.load(f)[name][f"{solver}/{jacobian}"]["flops"], "op-counting oracle build (tests/golden/flops.json)"
    except (OSError, KeyError):
        return attempt_flops(mech, solver), "hand model (fcs/flops.py)"


def host_threads():
    """Host cores this process may use, at most 16 (one GPU's CPU share)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def error_vs_oracle(mech, field, idx, runs):
    """Error of GPU results (runs: name -> download() dict over the whole
    field) on the cells idx against the oracle's relTol 1e-10 / absTol 1e-14
    solve of the same cells, per cell class, with the oracle's own relTol
    1e-4 solve beside them.  Cells the tight solve fails on are left out."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    o = Oracle(mech)
    sub = field.subset(idx)
    th = host_threads()
    t0 = time.perf_counter()
    tight = o.solve(sub, DT, threads=th, Treact=TREACT, absTol=1e-14, relTol=1e-10, maxSteps=2000000)
    t_tight = time.perf_counter() - t0
    ok = (tight["status"] & 6) == 0
    rho = tight["rho0"]
    Yt = sub.Y + tight["RR"] * DT / rho[None, :]
    ref = o.solve(sub, DT, threads=th, Treact=TREACT)
    rows = {"oracle_relTol_1e-4": {k: v for k, v in ref.items()}}
    for name, r in runs.items():
        rows[name] = {"RR": r["RR"][:, idx], "Tend": r["Tend"][idx]}
    out = {}
    for name, r in rows.items():
        Y = sub.Y + r["RR"] * DT / rho[None, :]
        per = {}
        for c, cname in ((1, "product"), (2, "flame_front")):
            m = ok & (sub.cls == c)
            dY = np.abs(Y - Yt)[:, m]
            Ym = np.abs(Yt[:, m])
            per[cname] = dict(max_abs_Y=float(dY.max()),
                              max_rel_Y=float((dY / np.maximum(Ym, 1e-6)).max()),
                              max_tol_scaled_Y=float((dY / (1e-8 + 1e-4 * Ym)).max()),
                              max_rel_T=float((np.abs(r["Tend"][m] - tight["Tend"][m]) / tight["Tend"][m]).max()))
        out[name] = per
    worst = lambda k: max(v[k] for v in out["learned"].values())
    out["sample"] = (f"{int((sub.cls == 1).sum())} product + {int((sub.cls == 2).sum())} flame-front cells "
                     f"(stratified, seed 3); reference: oracle OptRodas34 at relTol 1e-10 / absTol 1e-14 "
                     f"({int(ok.sum())} of {idx.size} cells integrated, {t_tight:.1f} s on {th} threads); "
                     "max_rel_Y = |dY| / max(|Y|, 1e-6), max_tol_scaled_Y = |dY| / (absTol + relTol |Y|)")
    out["headline_bar"] = dict(max_abs_Y=worst("max_abs_Y"), max_rel_T=worst("max_rel_T"),
                               holds=bool(worst("max_abs_Y") <= 1e-4 and worst("max_rel_T") <= 1e-4),
                               reading="|dY| <= 1e-4 of the unit mass and |dT| <= 1e-4 T, learned run")
    return out


def timed_runs(batch, steps, warmup, profile, label=None):
    """label: log a line after every step (runs of a minute per step)."""
    for k in range(warmup):
        t0 = time.perf_counter()
        batch.run(DT, timed=True)
        if label:
            log(f"{label}: warm-up step {k + 1} of {warmup}, {time.perf_counter() - t0:.1f} s")
    batch.profile(profile)
    launches, step_ms = 0, 0.0
    t0 = time.perf_counter()
    for k in range(steps):
        batch.run(DT, timed=True)
        n, ms = batch.stats()
        launches += n
        step_ms += ms
        if label:
            log(f"{label}: step {k + 1} of {steps} done")
    wall = time.perf_counter() - t0
    batch.profile(False)
    return wall, launches, step_ms


def stratified(field, n_per, seed=3):
    rng = np.random.default_rng(seed)
    idx = [rng.choice(np.nonzero(field.cls == c)[0], n_per, replace=False) for c in (1, 2)]
    return np.concatenate(idx)


def attempt(mech, gm):
    """Two launches of one Rodas34 attempt (dx = 1e-9 s) on the first 1,024
    cells of the mixture-fraction field: one wavefront per cell, the unit of
    work of the step kernel, for hardware-counter collection."""
    from fcs import lib
    from fcs.mechanism import config_field
    fld = config_field(1, mech, n_cells=1024)
    s = lib.Solver(gm, solver="OptRodas34")
    YT = np.vstack([fld.Y, fld.T[None, :]])
    t0 = time.perf_counter()
    for _ in range(2):
        err, _, _ = s.attempt(fld.p, YT, np.full(fld.n_cells, 1e-9))
    print(json.dumps({"workload": "2 x one Rodas34 attempt on 1,024 cells", "seconds": time.perf_counter() - t0,
                      "mean_error_norm": float(err.mean())}), flush=True)


def traffic(mech, gm):
    """The default workload's plain step with eager launches (no hipGraph):
    the 262,144-cell counterflow field, one plain Rodas34 step of
    k_step<56,Rodas34> launches after a k_setup, for HBM-byte counter
    collection (FETCH_SIZE, WRITE_SIZE per dispatch)."""
    from fcs import lib
    from fcs.mechanism import config_field
    solver = lib.Solver(gm, solver="OptRodas34", jacobian="exact", absTol=1e-8, relTol=1e-4, Treact=TREACT)
    field = config_field(2, mech, equilibrate=lib.equilibrator(solver))
    b = lib.Batch(solver, field)
    b.set_launch(0, 0)
    _, launches, step_ms = timed_runs(b, 1, 0, profile=True)
    r = b.download()
    print(json.dumps({"workload": "plain OptRodas34 step, eager launches, 262,144-cell counterflow field",
                      "launches": launches, "attempts": int(r["nAttempts"].sum()),
                      "avg_launch_ms": step_ms / max(launches, 1)}), flush=True)


def copy_calibration():
    """Three device-to-device copies of 256 MiB (the runtime's copy kernel)
    and three elementwise negations of 256 MiB: known-size references for
    the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 256 << 20
    a = torch.ones(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    x = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.neg(x, out=y)          # an elementwise kernel: reads and writes 256 MiB
    torch.cuda.synchronize()
    print(json.dumps({"workload": "3 x 256 MiB device copy, 3 x 256 MiB elementwise negation",
                      "bytes_per_launch": n}), flush=True)


def tutorials():
    """Plain Rodas34 (exact Jacobian, absTol 1e-8, relTol 1e-4) on the
    reference tutorials' mechanisms of tests/golden/tutorial_mechs.npz: each
    case's states tiled to 16,384 cells, deltaT 1e-5 s, best of two timed
    solves after a warm-up.  One JSON line per mechanism."""
    from fcs import lib
    from fcs.mechanism import Field, Mechanism
    g = np.load(os.path.join(ROOT, "tests", "golden", "tutorial_mechs.npz"))
    names = sorted({k.split("__")[0] for k in g.files}, key=lambda n: int(g[n + "__mech_n_species"]))
    for name in names:
        pre = name + "__mech_"
        mech = Mechanism.from_arrays({k[len(pre):]: g[k] for k in g.files if k.startswith(pre)})
        T, p, Y = g[name + "__T"], g[name + "__p"], g[name + "__Y"]
        rep = 16384 // T.size
        field = Field(np.tile(T, rep), np.tile(p, rep), np.tile(Y, (1, rep)), np.full(T.size * rep, 1e-7))
        b = lib.Batch(lib.Solver(lib.Mech(mech), solver="OptRodas34"), field)
        b.run(1e-5)
        best = min(_timed_run(b) for _ in range(2))
        att = int(b.download()["nAttempts"].sum())
        print(json.dumps({"mechanism": name, "species": mech.n_species, "reactions": mech.n_reactions,
                          "cells": field.n_cells, "cells_dt_per_s": field.n_cells / best,
                          "attempts_per_s": att / best}), flush=True)


PHASES = ["jacobian: thermo + reaction round", "jacobian: sparse term walk", "jacobian: production + dq/dT walk",
          "jacobian: third-body partials", "jacobian: dC/dY conversion + T row", "iteration matrix",
          "LU factorisation", "stage solves (all)", "RHS: density + thermo", "RHS: third-body sums",
          "RHS: reaction round", "RHS: production rates", "RHS: T equation", "stage sums", "error norm"]
LU_PARTS = ["LU: panels", "LU: U rows", "LU: trailing updates"]


def phases(args):
    """Phase split of the step kernel on the default field (plain Rodas34,
    eager launches) from the profiling build libfcs_phases.so: shader-clock
    cycles per phase, summed over cells by each cell's first thread, divided
    by the attempts made.  Build it first with
    `python fastchemistrysolver-openfoam-10_amd/fcs/build.py --phase-timing`."""
    import ctypes as C
    from fcs import build, lib
    from fcs.mechanism import config_field, config_mechanism
    L = lib.load(build.LIB_PHASES)
    L.fcs_phase_read.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.c_int]
    cfg = args.phase_config   # 2: the default 53-species field; 4: the 160-species one (workgroup form)
    mech = config_mechanism(cfg)
    gm = lib.Mech(mech)
    solver = lib.Solver(gm, solver="OptRodas34", jacobian="exact", absTol=1e-8, relTol=1e-4, Treact=TREACT)
    field = config_field(cfg, mech, n_cells=args.phase_side ** 2, equilibrate=lib.equilibrator(solver))
    b = lib.Batch(solver, field)
    b.run(DT)
    buf = (C.c_uint64 * 24)()
    L.fcs_phase_read(gm._h, buf, 1)          # clear
    t0 = time.perf_counter()
    b.run(DT)
    wall = time.perf_counter() - t0
    L.fcs_phase_read(gm._h, buf, 0)
    att = int(b.download()["nAttempts"].sum())
    cyc = [buf[i] / att for i in range(len(PHASES))]
    tot = sum(cyc)
    kern = "k_step<56,Rodas34>" if cfg != 4 else "k_step<176,Rodas34> (workgroup form)"
    print(json.dumps({"workload": f"phase split of {kern}, plain Rodas34 on a {args.phase_side}x"
                                  f"{args.phase_side} counterflow field of the {mech.n_species}-species mechanism "
                                  f"(profiling build)",
                      "attempts": att, "wall_s": wall, "cycles_per_attempt": tot,
                      "phases": {n: {"cycles": c, "share": c / tot} for n, c in zip(PHASES, cyc)},
                      "lu_parts": {n: {"cycles": buf[15 + i] / att, "share": buf[15 + i] / att / tot}
                                   for i, n in enumerate(LU_PARTS)}}), flush=True)


def _timed_run(batch):
    t0 = time.perf_counter()
    batch.run(1e-5)
    return time.perf_counter() - t0


def measured_traffic():
    """HBM bytes per k_step<56,Rodas34> launch of the default workload, as
    measured by two counter runs and committed with the profiles (None when
    the summary is absent)."""
    import re
    path = os.path.join(ROOT, "profiles", "traffic_k_step_rodas34_53sp_counterflow_262144cells.md")
    try:
        with open(path) as f:
            m = re.search(r"hbm_bytes_per_launch\s*=\s*([0-9.eE+]+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def mixture(args, mech, gm):
    """Plain Rodas34 on the 65,536-cell 1-D mixture-fraction field, with the
    step kernel's roofline from an event pair around every step launch."""
    from fcs import lib
    from fcs.mechanism import config_field
    f1 = config_field(1, mech)
    s1 = lib.Solver(gm, solver="OptRodas34", jacobian="exact", absTol=1e-8, relTol=1e-4)
    b1 = lib.Batch(s1, f1)
    wall, launches, step_ms = timed_runs(b1, args.steps, args.warmup, profile=True)
    r = b1.download()
    fpa, fsrc = counted_flops(mech, "OptRodas34")
    att = int(r["nAttempts"].sum())
    avg_launch_s = 1e-3 * step_ms / max(launches, 1)
    achieved = fpa * att * args.steps / max(launches, 1) / avg_launch_s / 1e12
    line = {"metric": METRIC, "value": f1.n_cells / (wall / args.steps), "unit": "cells·dt/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "plain OptRodas34, 65,536-cell 1-D mixture-fraction field, 53-species/"
                                   "325-reaction synthetic mechanism, deltaT 1e-6 s", "cells": f1.n_cells,
                       "attempts_per_step": att, "failed_cells": int(((r["status"] & 6) != 0).sum())},
            "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                         "kernel": "k_step<56,Rodas34>", "flops_per_attempt": fpa, "flops_source": fsrc, "launches": launches,
                         "avg_launch_ms": 1e3 * avg_launch_s}}
    print(json.dumps(line), flush=True)


def large_kernels():
    """The workgroup form's components through the kernel-level entries, on
    2,048 random states of the 160-species mechanism (two launches each):
    derivatives, Jacobian, LU factor + one solve, one Rodas34 attempt --
    for per-kernel timing under a profiler."""
    from fcs import lib
    from fcs.mechanism import config_mechanism, uniform_field
    mech = config_mechanism(4)
    s = lib.Solver(lib.Mech(mech), solver="OptRodas34")
    n, ns1 = 2048, mech.n_species + 1
    fld = uniform_field(mech, n, seed=3)
    YT = np.vstack([fld.Y, fld.T[None, :]])
    A = np.random.default_rng(1).standard_normal((n, ns1, ns1)) + 4 * np.eye(ns1)
    t0 = time.perf_counter()
    for _ in range(2):
        s.derivatives(fld.p, YT)
        s.jacobian(fld.p, YT)
        s.lu_solve(A, np.ones((n, ns1)))
        s.attempt(fld.p, YT, np.full(n, 1e-9))
    print(json.dumps({"workload": "160-species kernel-level entries, 2,048 states, 2 launches each",
                      "seconds": time.perf_counter() - t0}), flush=True)


def large(args):
    """Configuration 4: the 160-species / 1,200-reaction mechanism over a
    512-column counterflow field (131,072 cells by default), one workgroup
    per cell, matrix slices in HBM, LU trailing updates on FP64 MFMA.
    `value` is the learned-seeded Rodas34 run (network 163-256-256-256-162
    trained first, untimed, on two 512 x 16 fields); plain Rodas34 with the
    exact and the fast Jacobian and plain Seulex run beside it.  The step
    kernel's roofline comes from an event pair around every step launch of
    the plain exact run."""
    from fcs import learned, lib
    from fcs.mechanism import config_field, config_mechanism, counterflow_field
    mech = config_mechanism(4)
    gm = lib.Mech(mech)
    exact = lib.Solver(gm, solver="OptRodas34", jacobian="exact", absTol=1e-8, relTol=1e-4, Treact=TREACT)
    eq = lib.equilibrator(exact)
    t0 = time.perf_counter()
    field = config_field(4, mech, n_cells=args.large_cells, equilibrate=eq)
    log(f"field: {field.n_cells} cells in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    Xs, Ts = zip(*(learned.samples(exact, counterflow_field(mech, 512, seed=sd, equilibrate=eq, n_rows=16), DT)
                   for sd in (4001, 4002)))
    X, T = np.concatenate(Xs), np.concatenate(Ts)
    net, hist = learned.train(X, T, hidden=(256, 256, 256), epochs=args.large_epochs, batch=512, lr=2e-3, seed=1)
    t_train = time.perf_counter() - t0
    log(f"training: {X.shape[0]} samples, loss {hist[0]:.3g} -> {hist[-1]:.3g}, {t_train:.1f} s")
    runs = {}
    for name, solver, jac, mlp in (("plain", "OptRodas34", "exact", None), ("learned", "OptRodas34", "exact", net),
                                   ("fast", "OptRodas34", "fast", None), ("seulex", "OptSeulex", "exact", None)):
        s = exact if (solver, jac) == ("OptRodas34", "exact") else \
            lib.Solver(gm, solver=solver, jacobian=jac, absTol=1e-8, relTol=1e-4, Treact=TREACT)
        b = lib.Batch(s, field)
        if mlp is not None:
            b.set_learned(mlp, "verify")
        wall, launches, step_ms = timed_runs(b, args.steps, args.warmup, profile=(name == "plain"), label=name)
        r = b.download()
        runs[name] = dict(wall=wall, launches=launches, step_ms=step_ms, att=int(r["nAttempts"].sum()),
                          failed=int(((r["status"] & 6) != 0).sum()),
                          value=field.n_cells / (wall / args.steps))
        if mlp is not None:
            runs[name]["stats"] = b.learned_stats()
        if name == "plain":
            runs[name]["res"] = r
        log(f"{name}: {wall / args.steps * 1e3:.1f} ms per step, {runs[name]['att']} attempts")
        del b
    pl, ln = runs["plain"], runs["learned"]
    fpa, fsrc = counted_flops(mech, "OptRodas34")
    avg_launch_s = 1e-3 * pl["step_ms"] / max(pl["launches"], 1)
    achieved = fpa * pl["att"] / max(pl["launches"], 1) / avg_launch_s / 1e12
    npred, nver, nacc = ln["stats"]
    line = {"metric": "cells·dt/s at 160-species mech (learned-seeded vs plain GPU Rodas34, workgroup-per-cell "
                      "LU on FP64 MFMA)",
            "value": ln["value"], "unit": "cells·dt/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * ln["wall"] / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"learned-seeded OptRodas34 (verify mode), jacobian exact, 512x"
                                   f"{field.n_cells // 512} synthetic counterflow field ({field.n_cells} cells), "
                                   "160-species/1,200-reaction synthetic mechanism, deltaT 1e-6 s, Treact 500 K",
                       "cells": field.n_cells, "network": "163-256-256-256-162 SiLU, BF16 MFMA / FP32 accumulation",
                       "training_samples": int(X.shape[0]), "training_s": t_train,
                       "proposals": npred, "verified": nver, "accepted": nacc,
                       "attempts_per_step": ln["att"] // args.steps, "failed_cells": ln["failed"]},
            "plain": {"value": pl["value"], "attempts_per_step": pl["att"] // args.steps, "failed_cells": pl["failed"],
                      "speedup_learned_vs_plain": ln["value"] / pl["value"]},
            "fast_jacobian": {"value": runs["fast"]["value"], "attempts_per_step": runs["fast"]["att"] // args.steps,
                              "failed_cells": runs["fast"]["failed"]},
            "seulex": {"value": runs["seulex"]["value"], "columns_per_step": runs["seulex"]["att"] // args.steps,
                       "failed_cells": runs["seulex"]["failed"]},
            "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                         "kernel": "k_step<176,Rodas34> (one workgroup of 256 threads per active cell), plain run",
                         "flops_per_attempt": fpa, "flops_source": fsrc, "launches": pl["launches"],
                         "avg_launch_ms": 1e3 * avg_launch_s}}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(mech, field, n_cells=args.cpu_cells_large, gpu=pl.get("res"))
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true", help="skip the 65,536-cell plain run")
    ap.add_argument("--train-side", type=int, default=256, help="side of each training field")
    ap.add_argument("--train-samples", type=int, default=2_000_000,
                    help="training samples of the default workload's network (whole fields, so a few more)")
    ap.add_argument("--epochs", type=int, default=5)
    ap.add_argument("--large-epochs", type=int, default=30, help="training epochs of the large workload's network")
    ap.add_argument("--workload", choices=("counterflow", "mixture", "attempt", "large", "large-kernels",
                                           "traffic", "copy", "tutorials", "phases"),
                    default="counterflow",
                    help="counterflow: learned-seeded vs plain Rodas34 on the 262,144-cell field (default); "
                         "mixture: plain Rodas34 on the 65,536-cell mixture-fraction field only; "
                         "attempt: two single-attempt launches on 1,024 cells (for counter collection); "
                         "large: plain Rodas34 on the 160-species mechanism's counterflow field; "
                         "large-kernels: its components through the kernel-level entries (for profiling); "
                         "traffic: one plain step of the default field with eager launches (for byte counters); "
                         "copy: 256 MiB device copies (the byte counters' calibration); "
                         "tutorials: plain Rodas34 on the reference tutorials' mechanisms; "
                         "phases: per-phase cycles of the step kernel (profiling build libfcs_phases.so)")
    ap.add_argument("--phase-side", type=int, default=128, help="side of the phases workload's field")
    ap.add_argument("--phase-config", type=int, default=2, choices=(2, 4),
                    help="mechanism of the phases workload (2: 53 species, 4: 160 species)")
    ap.add_argument("--large-cells", type=int, default=131072, help="cells of the large workload (x 512)")
    ap.add_argument("--cpu-cells", type=int, default=4096,
                    help="cells of the CPU baseline's stratified sample (headline field)")
    ap.add_argument("--cpu-cells-large", type=int, default=512,
                    help="cells of the CPU baseline's sample on the 160-species workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); cells are independent, "
              "so the path shards by cell ranges with no collective on the data path.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "large":
        return large(args)
    if args.workload == "large-kernels":
        return large_kernels()
    if args.workload == "copy":
        return copy_calibration()
    if args.workload == "tutorials":
        return tutorials()
    if args.workload == "phases":
        return phases(args)
    from fcs import learned, lib
    from fcs.mechanism import config_field, config_mechanism, counterflow_field

    mech = config_mechanism(2)
    gm = lib.Mech(mech)
    if args.workload == "mixture":
        return mixture(args, mech, gm)
    if args.workload == "attempt":
        return attempt(mech, gm)
    if args.workload == "traffic":
        return traffic(mech, gm)
    solver = lib.Solver(gm, solver="OptRodas34", jacobian="exact", absTol=1e-8, relTol=1e-4, Treact=TREACT)
    eq = lib.equilibrator(solver)
    t0 = time.perf_counter()
    field = config_field(2, mech, equilibrate=eq)
    log(f"field: {field.n_cells} cells in {time.perf_counter() - t0:.1f} s")

    # ---- learned seed: training data from the plain integrator, training (untimed)
    t0 = time.perf_counter()
    Xs, Ts = [], []
    seed, n_samples = 2001, 0
    while n_samples < args.train_samples:   # seeded fields until the sample count is reached
        f = counterflow_field(mech, args.train_side, seed=seed, equilibrate=eq)
        X, T = learned.samples(solver, f, DT)
        Xs.append(X)
        Ts.append(T)
        n_samples += X.shape[0]
        seed += 1
    X, T = np.concatenate(Xs), np.concatenate(Ts)
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    net, hist = learned.train(X, T, epochs=args.epochs, batch=512, lr=2e-3, seed=1)
    t_train = time.perf_counter() - t0
    log(f"training: {X.shape[0]} samples in {t_data:.1f} s, {args.epochs} epochs in {t_train:.1f} s, "
        f"loss {hist[0]:.3g} -> {hist[-1]:.3g}")

    fpa, fsrc = counted_flops(mech, "OptRodas34")

    # ---- plain Rodas34 on the field
    plain = lib.Batch(solver, field)
    plain.set_launch(0, GRAPH_LAUNCHES)
    wall_p, _, _ = timed_runs(plain, args.steps, args.warmup, profile=False)
    rp = plain.download()
    log(f"plain: {wall_p / args.steps * 1e3:.1f} ms per step")

    # ---- the same plain step through fcs_solve over host arrays (the
    # OpenFOAM call's boundary: per-species host pointers), so the rate
    # includes the uploads and downloads over PCIe; reported beside value
    solver.solve(field, DT)
    t0 = time.perf_counter()
    rh = solver.solve(field, DT)
    wall_h = time.perf_counter() - t0
    host = {"api": "fcs_solve over host arrays (per-species pointers), plain OptRodas34, one flow step",
            "value": field.n_cells / wall_h, "ms_per_step": 1e3 * wall_h,
            "RR_bit_identical_to_batch": bool(np.array_equal(rh["RR"], rp["RR"]))}
    log(f"host arrays: {wall_h * 1e3:.1f} ms per step")

    # ---- learned-seeded Rodas34 on the same field (the measured workload)
    lrn = lib.Batch(solver, field)
    lrn.set_launch(0, GRAPH_LAUNCHES)
    lrn.set_learned(net, "verify")
    wall_l, _, _ = timed_runs(lrn, args.steps, args.warmup, profile=False)
    rl = lrn.download()
    npred, nver, nacc = lrn.learned_stats()
    log(f"learned: {wall_l / args.steps * 1e3:.1f} ms per step, {nacc}/{npred} proposals accepted")
    # Amdahl bound of the learned seed on this field: the attempts the plain
    # run spends on the cells the gate admitted (accepted proposals) and on
    # all product cells, against the plain run's total; the speed-up if every
    # such cell took one attempt (plus one verification per rejected proposal)
    acc = (rl["status"] & 8) != 0
    att_total = int(rp["nAttempts"].sum())
    att_acc = int(rp["nAttempts"][acc].sum())
    att_prod = int(rp["nAttempts"][field.cls == 1].sum())
    amdahl = {"attempt_share_of_accepted_cells": att_acc / att_total,
              "attempt_share_of_product_cells": att_prod / att_total,
              "attempt_share_of_flame_front_cells": int(rp["nAttempts"][field.cls == 2].sum()) / att_total,
              "max_speedup_admitted_cells_in_one_attempt": att_total / (att_total - att_acc + nacc + (nver - nacc)),
              "max_speedup_every_product_cell_in_one_attempt":
                  att_total / (att_total - att_prod + int((field.cls == 1).sum()))}

    # ---- seed mode: proposals verified as above, and every other reacting
    # cell starts from the network's first sub-step min(h, deltaT) instead of
    # its deltaTChem (so its results differ from the plain run's within the
    # tolerance); reported beside value
    sdb = lib.Batch(solver, field)
    sdb.set_launch(0, GRAPH_LAUNCHES)
    sdb.set_learned(net, "seed")
    wall_s, _, _ = timed_runs(sdb, args.steps, args.warmup, profile=False)
    rs = sdb.download()
    log(f"learned, seed mode: {wall_s / args.steps * 1e3:.1f} ms per step")

    # roofline of the step kernel on the learned run: an event pair around
    # every eager step launch on the solver's stream
    lrn.set_launch(0, 0)
    _, launches, step_ms = timed_runs(lrn, 1, 0, profile=True)
    list_len, grid = lrn.load_report()   # the same profiled run: one list length per step launch
    att_steps = int(rl["nAttempts"][(rl["status"] & 8) == 0].sum())
    avg_launch_s = 1e-3 * step_ms / max(launches, 1)
    achieved = fpa * att_steps / max(launches, 1) / avg_launch_s / 1e12

    # ---- plain Rodas34 on the 65,536-cell mixture-fraction field
    secondary = None
    if not args.no_secondary:
        f1 = config_field(1, mech)
        s1 = lib.Solver(gm, solver="OptRodas34", jacobian="exact", absTol=1e-8, relTol=1e-4)
        b1 = lib.Batch(s1, f1)
        w1, _, _ = timed_runs(b1, 1, 1, profile=False)
        secondary = {"workload": "plain OptRodas34, 65,536-cell 1-D mixture-fraction field, deltaT 1e-6 s",
                     "value": f1.n_cells / w1, "ms_per_step": 1e3 * w1,
                     "attempts_per_step": int(b1.download()["nAttempts"].sum())}

    value = field.n_cells / (wall_l / args.steps)
    plain_value = field.n_cells / (wall_p / args.steps)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cells·dt/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall_l / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "learned-seeded OptRodas34 (verify mode), 512x512 synthetic counterflow field "
                               "(262,144 cells), 53-species/325-reaction synthetic mechanism, deltaT 1e-6 s, "
                               f"Treact 500 K, step launches replayed from {GRAPH_LAUNCHES}-launch hipGraphs",
                   "cells": field.n_cells, "species": mech.n_species, "reactions": mech.n_reactions,
                   "solver": "OptRodas34", "jacobian": "exact", "absTol": 1e-8, "relTol": 1e-4,
                   "network": "56-128-128-128-55 SiLU, BF16 MFMA / FP32 accumulation",
                   "training_samples": int(X.shape[0]), "training_fields": len(Xs), "training_epochs": args.epochs,
                   "training_s": t_data + t_train},
        "plain": {"value": plain_value, "ms_per_step": 1e3 * wall_p / args.steps,
                  "attempts_per_step": int(rp["nAttempts"].sum())},
        "learned": {"speedup_vs_plain": value / plain_value, "proposals": npred, "verified": nver,
                    "accepted": nacc, "acceptance_rate": nacc / max(npred, 1),
                    "attempts_per_step": int(rl["nAttempts"].sum()),
                    "fallback_cells_bit_identical": bool(all(
                        np.array_equal(rl[k][..., (rl["status"] & 8) == 0], rp[k][..., (rl["status"] & 8) == 0])
                        for k in ("RR", "Tend", "nAttempts"))),
                    "failed_cells": int(((rl["status"] & 6) != 0).sum()),
                    "amdahl_bound": amdahl},
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": measured_traffic(),
                     "traffic_unit": "HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, copy-calibrated; "
                                     "profiles/traffic_k_step_rodas34_53sp_counterflow_262144cells.md)",
                     "kernel": "k_step<56,Rodas34> (FP64 VALU; one Rosenbrock attempt per active cell)",
                     "flops_per_attempt": fpa, "flops_source": fsrc, "launches": launches,
                     "avg_launch_ms": 1e3 * avg_launch_s},
        "load_per_wavefront": dict(lib.load_summary(list_len, grid),
                                   note="step launches of the roofline run (learned batch, eager launches); "
                                        "a launch strides its active list over the grid's wavefronts"),
    }
    line["host_buffers"] = host
    line["learned_seed_mode"] = {"value": field.n_cells / (wall_s / args.steps),
                                 "speedup_vs_plain": (field.n_cells / (wall_s / args.steps)) / plain_value,
                                 "attempts_per_step": int(rs["nAttempts"].sum()),
                                 "failed_cells": int(((rs["status"] & 6) != 0).sum())}
    if secondary:
        line["plain_mixture_fraction_65536"] = secondary
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(mech, field, n_cells=args.cpu_cells, gpu=rp)
        line["error_vs_oracle_tight"] = error_vs_oracle(mech, field, stratified(field, 48),
                                                        {"plain": rp, "learned": rl, "learned_seed_mode": rs})
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
