"""The step kernel's source run on the CPU with one thread per lane of a
64-lane wavefront (tests/emu): every cross-lane operation (readlane,
readfirstlane, DPP, shuffle, ballot) and every barrier is a rendezvous of
all 64 lanes, and each one checks that the lanes arrived from the same call
site with the same lane argument -- control flow that feeds a cross-lane
operation must be wave-uniform for the hardware's result to be the one the
source means.  The emulated k_step runs whole solve(deltaT) calls over the
library's own mechanism tables (fcs_mech_tables.h) and must match the
oracle: the same attempts per cell, RR and T at round-off.  The learned
seed's gate and k_verify run in front of the step launches the way the
library orders them, on proposals made from the oracle's own Rodas34 step.

Bar: no rendezvous with lanes at different call sites or arguments, none
left incomplete; per cell the oracle's attempt count and status; RR within
1e-10 of the largest |RR| relative, T within 1e-8 K."""
import numpy as np
import pytest

import fcs.mechanism as M
from oracle import Oracle

from emu import emu

CELL_LEARNED = 8


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


def oracle_proposals(orc, field, dt, hs, zpert, absTol=1e-8, relTol=1e-4):
    """Proposal rows [ln h, z] whose end state is the oracle's Rodas34 step
    over dt from each cell's clipped state, compressed the way k_verify reads
    them (d = sinh(z) (absTol + relTol |y0|)), plus zpert; h = hs dt."""
    ns = orc.mech.n_species
    P = np.zeros((field.n_cells, ns + 2), np.float32)
    for c in range(field.n_cells):
        YT = np.concatenate([np.maximum(field.Y[:, c], 0), [field.T[c]]])
        err, yT, y0 = orc.attempt(field.p[c], YT, dt)
        P[c, 0] = np.log(hs[c] * dt)
        P[c, 1:] = np.arcsinh((yT - y0) / (absTol + relTol * np.abs(y0))) + zpert[c]
    return P


def equilibrate(orc, field, cells, t):
    for c in cells:
        YT = np.concatenate([np.maximum(field.Y[:, c], 0), [field.T[c]]])
        r = orc.integrate_cell(field.p[c], YT, t, 1e-7)
        field.Y[:, c] = np.maximum(r["y"][:-1], 0)
        field.T[c] = r["y"][-1]


def assert_emulated_learned(mech, field, dt, hs, zpert, accepted):
    """Seed mode: accepted cells finish on the verified step (status
    CELL_LEARNED, one attempt), every other cell integrates from
    min(h, dt); the oracle's solve from those first sub-steps is the
    reference for all of them (for an accepted cell it takes the same single
    step over dt)."""
    orc = Oracle(mech)
    P = oracle_proposals(orc, field, dt, hs, zpert)
    r = emu.solve(mech, field, dt, P=P, seed=1)
    assert r["violations"] == 0, r["message"]
    assert r["rc"] == 0 and r["checks"] > 1000
    gated = np.exp(P[:, 0].astype(np.float64)) >= 0.5 * dt
    assert r["nAccepted"] == (int(accepted.sum()), int(gated.sum()))
    assert ((r["status"] & CELL_LEARNED != 0) == accepted).all()
    fs = field.subset(np.arange(field.n_cells))
    fs.deltaTChem = np.minimum(np.exp(P[:, 0].astype(np.float64)), dt)
    o = orc.solve(fs, dt)
    assert (r["status"] & ~CELL_LEARNED == o["status"]).all()
    assert (o["nAttempts"][accepted] == 1).all()
    assert (r["nAttempts"] == o["nAttempts"]).all(), (r["nAttempts"], o["nAttempts"])
    assert np.abs(r["RR"] - o["RR"]).max() <= 1e-10 * np.abs(o["RR"]).max()
    assert np.abs(r["Tend"] - o["Tend"]).max() <= 1e-8
    # the step proposal follows the error norm's fourth root; the norm of a
    # near-equilibrium step is a difference of nearly equal numbers
    np.testing.assert_allclose(r["deltaTChem"], o["deltaTChem"], rtol=1e-6)



def test_emulated_learned_verify_ten_species(mech0):
    """Gate and k_verify<12> on four cells, three of them near equilibrium
    after 1e-3 s of integration: an exact proposal (accepted), one with a
    species increment 4 units of z off (fails the agreement test), one with
    a NaN increment, and a flame cell whose Rodas34 step over dt fails its
    own error test; then the step launches for the three rejected cells."""
    field = M.config_field(0, mech0, n_cells=4)
    equilibrate(Oracle(mech0), field, (0, 1, 2), 1e-3)
    zp = np.zeros((4, mech0.n_species + 1))
    zp[1, 3] = 4.0
    zp[2, 2] = np.nan
    assert_emulated_learned(mech0, field, 1e-7, np.array([1.5, 1.5, 2.0, 1.5]), zp,
                            np.array([True, False, False, False]))


def test_emulated_learned_verify_gri30(gri30):
    """Gate and k_verify<56> (the benchmark's mechanism) on three hot
    snapshot cells brought near equilibrium: an exact proposal (accepted), a
    perturbed one (rejected, then integrated from h = deltaT), and one with
    h = deltaT/10 (not gated; seeded)."""
    mech, field = gri30
    f = field.subset(np.argsort(field.T)[-3:])
    equilibrate(Oracle(mech), f, (0, 1, 2), 1e-4)
    zp = np.zeros((3, mech.n_species + 1))
    zp[1, 7] = -3.0
    assert_emulated_learned(mech, f, 1e-6, np.array([1.5, 1.5, 0.1]), zp, np.array([True, False, False]))
