"""CPU check of bench.py's cpu_baseline leg: the oracle's AVX2 build timed
on one core and then on every usable host core over the same stratified
sample, on the 10-species mechanism with a tiny three-class field; the
comparison block is fed the strict oracle's results in place of a GPU run."""
import os
import sys

import numpy as np

import fcs.mechanism as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from oracle import Oracle  # noqa: E402


def _field(mech):
    field = M.uniform_field(mech, 24, seed=5)
    field.cls = (np.arange(field.n_cells) % 3).astype(np.int32)
    return field


def test_cpu_baseline_reports_one_core_and_all_cores():
    sys.path.insert(0, ROOT)
    import bench
    mech = M.config_mechanism(0)
    field = _field(mech)
    strict = Oracle(mech).solve(field, bench.DT, Treact=bench.TREACT)
    r = bench.cpu_baseline(mech, field, n_cells=12, gpu=strict, runs=2)
    assert r["cores"] == 1 and r["kind"] == "port" and r["unit"] == "cells·dt/s"
    assert np.isfinite(r["value"]) and r["value"] > 0
    assert "12 cells" in r["sample"] and "-mavx2 -mfma" in r["sample"]
    assert r["host"]["nproc"] >= 1 and r["host"]["model"]
    ac = r["all_cores"]
    assert 1 <= ac["cores"] <= 16
    assert np.isfinite(ac["value"]) and ac["value"] > 0
    assert f"on {ac['cores']} threads" in ac["sample"]
    # the AVX2 build is the same algorithm: it agrees with the strict build
    # far inside the CPU/GPU bar
    v = r["vs_gpu"]
    assert v["max_abs_dY"] < 1e-9 and v["max_abs_dT_K"] < 1e-6


def test_avx2_build_matches_strict_oracle():
    mech = M.config_mechanism(0)
    field = _field(mech)
    a = Oracle(mech).solve(field, 1e-5)
    b = Oracle(mech, variant="avx2").solve(field, 1e-5)
    assert np.array_equal(a["status"], b["status"])
    np.testing.assert_allclose(b["Tend"], a["Tend"], rtol=1e-10)
    np.testing.assert_allclose(b["RR"], a["RR"], rtol=1e-6, atol=1e-9 * np.abs(a["RR"]).max())
