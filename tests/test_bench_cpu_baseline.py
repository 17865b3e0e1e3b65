"""CPU check of bench.py's cpu_baseline leg: the oracle timed on one core
and then on every usable host core over the same stratified sample, on the
10-species mechanism with a tiny three-class field."""
import os
import sys

import numpy as np

import fcs.mechanism as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpu_baseline_reports_one_core_and_all_cores():
    sys.path.insert(0, ROOT)
    import bench
    mech = M.config_mechanism(0)
    field = M.uniform_field(mech, 24, seed=5)
    field.cls = (np.arange(field.n_cells) % 3).astype(np.int32)
    r = bench.cpu_baseline(mech, field, seconds=0.2)
    assert r["cores"] == 1 and r["kind"] == "port" and r["unit"] == "cells·dt/s"
    assert np.isfinite(r["value"]) and r["value"] > 0
    ac = r["all_cores"]
    assert 1 <= ac["cores"] <= 16
    assert np.isfinite(ac["value"]) and ac["value"] > 0
    assert f"on {ac['cores']} threads" in ac["sample"]
