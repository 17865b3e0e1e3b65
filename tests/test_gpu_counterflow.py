"""The benchmark's counterflow field on the GPU against the oracle.

Configuration 2 (plain OptRodas34 over a shuffled counterflow field of cold,
near-equilibrium product and flame-front cells, step launches replayed from
a hipGraph) and configuration 3 (the same field with the learned seed, in
verify and in seed mode):

* every cell of a 32 x 32 field against the oracle's solve(deltaT)
  (FastChemistryModel_transientSolve.H:35-83 restated in oracle/), plain,
  learned-accepted, learned-fallback and seed-mode cells alike;
* the full 512 x 512 field (262,144 cells) through size-independent
  properties -- statuses, cold-cell write-out, per-cell mass and element
  conservation, bit-identical repeats, graph replay equal to eager
  launches, learned fallback cells bit-identical to the plain run -- and a
  stratified sample of it against the oracle;
* the error of the plain, learned and seed-mode GPU results against a
  relTol 1e-10 / absTol 1e-14 oracle solve, per cell class, beside the
  error of the oracle's own relTol 1e-4 solve of the same cells.
"""
import os

import numpy as np
import pytest

import fcs.mechanism as M
from oracle import Oracle
from parity import assert_solve_parity, assert_state_parity, end_state, pick

pytestmark = pytest.mark.gpu

DT = 1e-6
TREACT = 500.0
CELL_LEARNED = 8
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def fcslib():
    from fcs import lib
    lib.load()
    assert lib.device_count() >= 1, "no GPU visible to the HIP runtime"
    return lib


@pytest.fixture(scope="module")
def env(fcslib, mech1):
    """Solver, equilibrator, the 32 x 32 field and a network trained on two
    other seeded 64 x 64 counterflow fields."""
    from fcs import learned
    gm = fcslib.Mech(mech1)
    s = fcslib.Solver(gm, Treact=TREACT)
    eq = fcslib.equilibrator(s)
    field = M.config_field(2, mech1, n_cells=32 * 32, equilibrate=eq)
    Xs, Ts = zip(*(learned.samples(s, M.counterflow_field(mech1, 64, seed=sd, equilibrate=eq), DT)
                   for sd in (31, 32)))
    net, hist = learned.train(np.concatenate(Xs), np.concatenate(Ts), epochs=40, batch=256, lr=2e-3, seed=1)
    assert hist[-1] < 0.5 * hist[0]
    return dict(gm=gm, s=s, eq=eq, field=field, net=net, oracle=Oracle(mech1))


def batch_run(fcslib, s, field, net=None, mode="verify", graph=24):
    b = fcslib.Batch(s, field)
    b.set_launch(0, graph)
    if net is not None:
        b.set_learned(net, mode)
    b.run(DT)
    r = b.download()
    stats = b.learned_stats()
    return r, stats


def oracle_solve(env, field, **opts):
    opts.setdefault("Treact", TREACT)
    return env["oracle"].solve(field, DT, threads=THREADS, **opts)


def seed_steps(fcslib, net, field):
    """The first sub-step seed mode gives each cell, min(h, deltaT) with h the
    network's proposal, from the same inputs the device builds
    (k_features: [T, ln p, max(0, Y), ln deltaT] in float32); NaN for cells
    whose proposal is not a positive number (they keep deltaTChem)."""
    X = np.concatenate([field.T[None, :], np.log(field.p)[None, :], np.maximum(field.Y, 0.0),
                        np.full((1, field.n_cells), np.log(DT))], axis=0).T.astype(np.float32)
    h = np.exp(net.forward(X)[:, 0].astype(np.float64))
    return np.where((h == h) & (h > 0), np.minimum(h, DT), np.nan)


# ---------------------------------------------------------- 32 x 32 field
def test_c2_plain_graph_replay_every_cell_vs_oracle(fcslib, env, mech1):
    f = env["field"]
    assert set(np.unique(f.cls)) == {0, 1, 2}
    g, _ = batch_run(fcslib, env["s"], f)
    o = oracle_solve(env, f)
    assert o["rc"] == 0 and (g["status"] & 6 == 0).all()
    g["rho0"] = o["rho0"]
    assert_solve_parity(g, o, f, DT, rerun=lambda f2: oracle_solve(env, f2))
    # flame-front cells carry the stiffness: many attempts, and the same
    # counts as the reference algorithm on nearly every cell
    flame = f.cls == 2
    assert flame.sum() >= 100 and g["nAttempts"][flame].min() >= 40
    assert np.mean(g["nAttempts"] == o["nAttempts"]) >= 0.95


def test_c3_learned_verify_every_cell_vs_oracle(fcslib, env, mech1):
    f = env["field"]
    plain, _ = batch_run(fcslib, env["s"], f)
    g, (npred, nver, nacc) = batch_run(fcslib, env["s"], f, env["net"], "verify")
    react = f.T > TREACT
    acc = (g["status"] & CELL_LEARNED) != 0
    assert npred == react.sum() and nacc == acc.sum() and nacc >= 0.1 * (f.cls == 1).sum()
    o = oracle_solve(env, f)
    rho0 = o["rho0"]
    # cells the network finished: one verified Rodas34 step over deltaT,
    # against the reference algorithm's full integration
    ga, oa = pick(g, acc), pick(o, acc)
    ga["status"] = ga["status"] & ~CELL_LEARNED
    assert_state_parity(ga, oa, f.subset(np.nonzero(acc)[0]), DT, rho0[acc], "learned-accepted cells")
    # every other cell: bit-identical to the plain run, and so to the oracle at the bar
    fb = ~acc
    for k in ("RR", "Tend", "deltaTChem", "nAttempts", "status"):
        np.testing.assert_array_equal(g[k][..., fb], plain[k][..., fb])
    assert_state_parity(pick(g, fb), pick(o, fb), f.subset(np.nonzero(fb)[0]), DT, rho0[fb], "fallback cells")


def test_c3_seed_mode_every_cell_vs_oracle(fcslib, env, mech1):
    """Seed mode: the cells the network does not finish start from its first
    sub-step; the oracle run from the same first sub-steps is the reference."""
    f = env["field"]
    g, (npred, nver, nacc) = batch_run(fcslib, env["s"], f, env["net"], "seed")
    react = f.T > TREACT
    acc = (g["status"] & CELL_LEARNED) != 0
    seeded = react & ~acc
    h = seed_steps(fcslib, env["net"], f)
    assert np.isfinite(h[seeded]).all()
    fs = f.subset(np.arange(f.n_cells))
    fs.deltaTChem[seeded] = h[seeded]
    o = oracle_solve(env, fs)
    idx = np.nonzero(seeded)[0]
    gs, os_ = pick(g, idx), pick(o, idx)
    assert_solve_parity(gs, os_, fs.subset(idx), DT, rho0=o["rho0"][idx], check_dmin=False,
                        rerun=lambda f2: oracle_solve(env, f2))
    # the seeds shorten the flame-front integrations' step search
    plain, _ = batch_run(fcslib, env["s"], f)
    assert g["nAttempts"][seeded].sum() <= 1.05 * plain["nAttempts"][seeded].sum()


# ---------------------------------------------------------- 512 x 512 field
@pytest.fixture(scope="module")
def full(fcslib, env, mech1):
    f = M.config_field(2, mech1, equilibrate=env["eq"])
    assert f.n_cells == 262144
    return f


def test_c2_c3_full_field_properties(fcslib, env, full, mech1):
    f = full
    s = env["s"]
    g1, _ = batch_run(fcslib, s, f)
    # statuses: cold cells skipped and written out unchanged, every other cell integrated
    cold = f.T <= TREACT
    assert np.abs(cold.mean() - 230 / 512) < 1e-12
    assert (g1["status"][cold] == 1).all() and (g1["status"][~cold] == 0).all()
    assert (g1["RR"][:, cold] == 0).all() and (g1["Tend"][cold] == f.T[cold]).all()
    np.testing.assert_array_equal(g1["deltaTChem"][cold], f.deltaTChem[cold])
    assert np.isfinite(g1["RR"]).all() and np.isfinite(g1["Tend"]).all()
    # per-cell conservation: the synthetic mechanism conserves mass and
    # four elements reaction by reaction
    rho0 = f.rho(mech1)
    Yend = end_state(g1, f, DT, rho0)
    assert np.abs(Yend.sum(axis=0) - f.Y.sum(axis=0)).max() <= 1e-9
    elem = mech1.composition.T @ (Yend / mech1.W[:, None]) - mech1.composition.T @ (f.Y / mech1.W[:, None])
    assert np.abs(elem).max() <= 1e-10
    # repeat (graph replay) and eager launches: bit-identical
    g2, _ = batch_run(fcslib, s, f)
    ge, _ = batch_run(fcslib, s, f, graph=0)
    for k in ("RR", "Tend", "deltaTChem", "nAttempts", "status"):
        np.testing.assert_array_equal(g1[k], g2[k])
        np.testing.assert_array_equal(g1[k], ge[k])
    assert g1["deltaTMin"] == g2["deltaTMin"] == ge["deltaTMin"]
    # learned seed at full size: fallback cells bit-identical to the plain run
    gl, (npred, nver, nacc) = batch_run(fcslib, s, f, env["net"], "verify")
    acc = (gl["status"] & CELL_LEARNED) != 0
    assert npred == (~cold).sum() and nacc == acc.sum() > 0
    for k in ("RR", "Tend", "deltaTChem", "nAttempts", "status"):
        np.testing.assert_array_equal(gl[k][..., ~acc], g1[k][..., ~acc])
    # a stratified sample of the full field against the oracle
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(np.nonzero(f.cls == c)[0], n, replace=False)
                          for c, n in ((0, 16), (1, 32), (2, 32))])
    sub = f.subset(idx)
    o = oracle_solve(env, sub)
    assert_state_parity(pick(g1, idx), o, sub, DT, rho0[idx], "plain sample")
    la = acc[idx]
    if la.any():
        gsub = pick(gl, idx[la])
        gsub["status"] = gsub["status"] & ~CELL_LEARNED
        assert_state_parity(gsub, pick(o, np.nonzero(la)[0]), sub.subset(np.nonzero(la)[0]), DT,
                            rho0[idx[la]], "learned sample")


# ---------------------------------------------------------- error vs a tight oracle solve
def test_error_against_tight_oracle(fcslib, env, mech1):
    """Error of the GPU results at relTol 1e-4 against the oracle at relTol
    1e-10 / absTol 1e-14 on 12 product and 12 flame-front cells, beside the
    oracle's own relTol 1e-4 error on the same cells: the GPU paths have the
    reference algorithm's error (the headline's "equal error")."""
    f = env["field"]
    rng = np.random.default_rng(11)
    idx = np.concatenate([rng.choice(np.nonzero(f.cls == c)[0], 12, replace=False) for c in (1, 2)])
    sub = f.subset(idx)
    tight = oracle_solve(env, sub, relTol=1e-10, absTol=1e-14, maxSteps=2000000)
    assert tight["rc"] == 0
    rho0 = tight["rho0"]
    Yt = end_state(tight, sub, DT)

    def err(r, cls):
        m = sub.cls == cls
        dY = np.abs(end_state(r, sub, DT, rho0) - Yt)[:, m]
        return dict(abs_Y=dY.max(), rel_Y=(dY / np.maximum(np.abs(Yt[:, m]), 1e-6)).max(),
                    tol_Y=(dY / (1e-8 + 1e-4 * np.abs(Yt[:, m]))).max(),
                    rel_T=(np.abs(r["Tend"] - tight["Tend"]) / tight["Tend"])[m].max())

    runs = {"oracle": oracle_solve(env, sub)}
    runs["plain"] = pick(batch_run(fcslib, env["s"], f)[0], idx)
    runs["verify"] = pick(batch_run(fcslib, env["s"], f, env["net"], "verify")[0], idx)
    runs["seed"] = pick(batch_run(fcslib, env["s"], f, env["net"], "seed")[0], idx)
    table = {k: {c: err(r, c) for c in (1, 2)} for k, r in runs.items()}
    for k, t in table.items():
        print(k, {c: {q: f"{v:.2e}" for q, v in e.items()} for c, e in t.items()})
    for c in (1, 2):
        ref = table["oracle"][c]
        for k in ("plain", "verify", "seed"):
            e = table[k][c]
            # same error as the reference algorithm at the same tolerances
            assert e["abs_Y"] <= 3.0 * ref["abs_Y"] + 1e-9, (k, c, e, ref)
            assert e["rel_T"] <= 3.0 * ref["rel_T"] + 1e-9, (k, c, e, ref)
