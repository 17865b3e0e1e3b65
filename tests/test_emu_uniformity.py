"""The step kernel's source run on the CPU with one thread per lane of a
64-lane wavefront (tests/emu): every cross-lane operation (readlane,
readfirstlane, DPP, shuffle, ballot) and every barrier is a rendezvous of
all 64 lanes, and each one checks that the lanes arrived from the same call
site with the same lane argument -- control flow that feeds a cross-lane
operation must be wave-uniform for the hardware's result to be the one the
source means.  The emulated k_step runs whole solve(deltaT) calls over the
library's own mechanism tables (fcs_mech_tables.h) and must match the
oracle: the same attempts per cell, RR and T at round-off.

Bar: no rendezvous with lanes at different call sites or arguments, none
left incomplete; per cell the oracle's attempt count and status; RR within
1e-10 of the largest |RR| relative, T within 1e-8 K."""
import numpy as np
import pytest

import fcs.mechanism as M
from oracle import Oracle

from emu import emu


def test_checker_reports_divergent_lanes():
    """The checker itself: a uniform read passes (and reads the right lane);
    lanes at two call sites, a lane-dependent lane argument, or a lane that
    skips the read are each reported."""
    r = emu.selftest(0)
    assert r["rc"] == 0 and r["violations"] == 0 and r["checks"] >= 2
    for kind, what in ((1, "call sites"), (2, "lane arguments"), (3, "call sites")):
        r = emu.selftest(kind)
        assert r["rc"] != 0 and r["violations"] >= 1, kind
        assert what in r["message"], (kind, r["message"])


def assert_emulated_solve(mech, field, dt, solver="OptRodas34"):
    r = emu.solve(mech, field, dt, solver=solver)
    assert r["violations"] == 0, r["message"]
    assert r["rc"] == 0 and r["checks"] > 10000
    o = Oracle(mech).solve(field, dt, solver=solver)
    assert (r["status"] == o["status"]).all()
    assert (r["nAttempts"] == o["nAttempts"]).all(), (r["nAttempts"], o["nAttempts"])
    assert np.abs(r["RR"] - o["RR"]).max() <= 1e-10 * np.abs(o["RR"]).max()
    assert np.abs(r["Tend"] - o["Tend"]).max() <= 1e-8
    return r


def test_emulated_step_kernel_ten_species(mech0):
    """Two cells of the 10-species mechanism through 1e-6 s (about 150 to
    200 attempts each: rejected and accepted steps, the last-step cut, cell
    completion), NP = 12."""
    field = M.config_field(0, mech0, n_cells=2)
    r = assert_emulated_solve(mech0, field, 1e-6)
    assert r["nAttempts"].min() > 100


def test_emulated_step_kernel_gri30(gri30):
    """Three flame cells of GRI-3.0 (53 species, 325 reactions: third-body,
    falloff and reverse rates, NP = 56) through 1e-6 s: the hottest cells of
    the converged snapshot, which take a few attempts each."""
    mech, field = gri30
    idx = np.argsort(field.T)[-3:]
    sub = field.subset(idx)
    r = assert_emulated_solve(mech, sub, 1e-6)
    assert r["nAttempts"].max() >= 1


@pytest.mark.parametrize("solver", ["OptRosenbrock34", "OptRodas23", "OptRosenbrock23", "OptSeulex"])
def test_emulated_step_kernels_other_solvers(mech0, solver):
    """The other four solvers' step kernels (k_step with their tableaus,
    k_step_seulex's extrapolation passes) on the 10-species mechanism
    through 1e-7 s: two cells, one for Seulex."""
    field = M.config_field(0, mech0, n_cells=1 if solver == "OptSeulex" else 2)
    assert_emulated_solve(mech0, field, 1e-7, solver)
