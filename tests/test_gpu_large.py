"""Parity of the workgroup form (systems of more than 64 rows: one
workgroup per cell, block-4-pivoted LU with FP64 MFMA trailing updates) with
the CPU oracle: the 160-species / 1,200-reaction synthetic mechanism (matrix
in HBM, 176 rows), an 80-species one (matrix in LDS, 96 rows, padding rows),
a 104-species one (matrix in LDS, 112 rows) and a 70-species one whose
5,000 reactions leave no LDS for the matrix (HBM, 128 rows)."""
import numpy as np
import pytest

import fcs.mechanism as M
from oracle import Oracle, lu_solve as orc_lu_solve

from test_gpu_parity import assert_solve_parity, fcslib  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mech4():
    return M.config_mechanism(4)


@pytest.fixture(scope="module")
def mech80():
    return M.synthetic_mechanism(80, 320, n_third_body=30, n_troe=10, n_plog=4, n_sri=3, n_noneq=4, seed=8080)


@pytest.fixture(scope="module")
def mech104():
    return M.synthetic_mechanism(104, 420, n_third_body=40, n_troe=12, n_plog=4, n_sri=2, n_noneq=6, seed=1104)


def yt_states(mech, n, seed):
    fld = M.uniform_field(mech, n, seed=seed)
    return fld, np.vstack([fld.Y, fld.T[None, :]])


@pytest.mark.parametrize("which", ["mech4", "mech80", "mech104"])
def test_large_lu_matches_oracle(fcslib, which, request):
    """Block-4-pivoted LU + solve of the workgroup form against the oracle's
    restatement of the reference LU, on random matrices that need pivoting
    (small diagonals), plus the signed-maximum pivot case of the reference."""
    mech = request.getfixturevalue(which)
    n = mech.n_species + 1
    s = fcslib.Solver(fcslib.Mech(mech))
    rng = np.random.default_rng(17)
    A = rng.standard_normal((6, n, n))
    A[:3] += 4.0 * np.eye(n)                         # well conditioned
    A[3:] *= rng.uniform(0.1, 10.0, (3, n, 1))       # row scaling: pivots move inside 4-blocks
    A[5, 1:4, 0] = 0.0
    A[5, 1, 1], A[5, 2, 1], A[5, 3, 1] = 1.0, -5.0, 2.0   # signed running maximum (LUsolver.C:177)
    b = rng.standard_normal((6, n))
    x = s.lu_solve(A, b)
    for k in range(6):
        xo, _, _ = orc_lu_solve(A[k], b[k])
        np.testing.assert_allclose(x[k], xo, rtol=1e-10, atol=1e-11 * np.abs(xo).max())
        assert np.abs(A[k] @ x[k] - b[k]).max() <= 1e-10 * np.abs(A[k]).max() * np.abs(x[k]).max()


@pytest.mark.parametrize("which", ["mech4", "mech80", "mech104"])
def test_large_derivatives_and_jacobian(fcslib, which, request):
    mech = request.getfixturevalue(which)
    o = Oracle(mech)
    fld, YT = yt_states(mech, 16, 3)
    YT[4, :3] = -1e-4
    for jac in ("exact", "fast"):
        s = fcslib.Solver(fcslib.Mech(mech), jacobian=jac)
        d, yc = s.derivatives(fld.p, YT)
        f, J, _ = s.jacobian(fld.p, YT)
        for c in range(16):
            dref, yref = o.derivatives(fld.p[c], YT[:, c])
            np.testing.assert_array_equal(yc[:, c], yref)
            scale = np.abs(dref[:-1]).max()
            assert np.abs(d[:-1, c] - dref[:-1]).max() <= 1e-11 * scale
            assert d[-1, c] == pytest.approx(dref[-1], rel=1e-10, abs=1e-11 * scale)
            fr, Jr, _ = o.jacobian(fld.p[c], YT[:, c], jac)
            np.testing.assert_allclose(f[:, c], fr, rtol=1e-10, atol=1e-11 * np.abs(fr[:-1]).max())
            rs = np.abs(Jr).max(axis=1, keepdims=True) + 1e-300
            assert (np.abs(J[c] - Jr) / rs).max() <= 1e-9


@pytest.mark.parametrize("solver", ["OptRodas34", "OptRosenbrock34", "OptRodas23"])
def test_large_single_attempt(fcslib, mech4, solver):
    o = Oracle(mech4)
    s = fcslib.Solver(fcslib.Mech(mech4), solver=solver)
    fld, YT = yt_states(mech4, 12, 4)
    dx = np.repeat([1e-13, 1e-11, 1e-9], 4)
    err, YTn, yc = s.attempt(fld.p, YT, dx)
    for c in range(12):
        er, ynr, yr = o.attempt(fld.p[c], YT[:, c], dx[c], solver=solver)
        np.testing.assert_array_equal(yc[:, c], yr)
        # The MFMA trailing update sums in another order than the reference,
        # and grossly rejected attempts amplify round-off.  Above
        # (0.9 / 0.2)^4 = 410 the controller's step scale is clamped at
        # minScale, so there only the magnitude of Err matters.
        assert err[c] == pytest.approx(er, rel=1e-4 if er < 410 else 1e-2, abs=1e-6), (c, err[c], er)
        inc = np.abs(ynr - yr).max()
        assert np.abs(YTn[:, c] - ynr).max() <= (1e-9 if er < 1.0 else 1e-3) * inc + 1e-15, c


@pytest.mark.parametrize("which,jac", [("mech4", "exact"), ("mech4", "fast"), ("mech80", "exact"),
                                       ("mech104", "exact"), ("mech104", "fast")])
def test_large_solve_matches_oracle(fcslib, which, jac, request):
    mech = request.getfixturevalue(which)
    fld = M.uniform_field(mech, 8, seed=21)
    dt = 1e-6
    g = fcslib.Solver(fcslib.Mech(mech), jacobian=jac).solve(fld, dt)
    o = Oracle(mech)
    r = o.solve(fld, dt, threads=1, jacobian=jac)
    assert g["rc"] == 0 and r["rc"] == 0
    assert_solve_parity(g, r, fld, dt, rerun=lambda f: o.solve(f, dt, threads=8, jacobian=jac))


def test_large_calculate_and_tc(fcslib, mech4):
    o = Oracle(mech4)
    fld = M.uniform_field(mech4, 16, seed=6)
    rho = fld.p / (M.RU * fld.T * (fld.Y / mech4.W[:, None]).sum(axis=0))
    s = fcslib.Solver(fcslib.Mech(mech4))
    RR = s.calculate(rho, fld.T, fld.p, fld.Y)
    tc = s.tc(rho, fld.T, fld.p, fld.Y)
    for c in range(16):
        rr = o.calculate(rho[c], fld.T[c], fld.p[c], fld.Y[:, c])
        np.testing.assert_allclose(RR[:, c], rr, rtol=1e-10, atol=1e-11 * np.abs(rr).max())
        assert tc[c] == pytest.approx(o.tc(rho[c], fld.T[c], fld.p[c], fld.Y[:, c]), rel=1e-10)


@pytest.mark.parametrize("which,n", [("mech4", 4), ("mech80", 8)])
def test_large_seulex_matches_oracle(fcslib, which, n, request):
    """OptSeulex in the workgroup form (per-cell Jacobian kept in HBM between
    steps, NP x NP doubles per cell)."""
    mech = request.getfixturevalue(which)
    fld = M.uniform_field(mech, n, seed=21)
    dt = 1e-6
    g = fcslib.Solver(fcslib.Mech(mech), solver="OptSeulex").solve(fld, dt)
    o = Oracle(mech)
    r = o.solve(fld, dt, threads=1, solver="OptSeulex")
    assert g["rc"] == 0 and r["rc"] == 0
    assert_solve_parity(g, r, fld, dt, rerun=lambda f: o.solve(f, dt, threads=8, solver="OptSeulex"))


def test_large_learned_seed(fcslib, mech4):
    """The learned seed on the 160-species mechanism (network 163 -> 256 x 3
    -> 162): fallback cells bit-identical to the plain integrator, accepted
    proposals within the parity bar of it, attempts saved on the
    near-equilibrium product cells."""
    from fcs import learned
    dt = 1e-6
    s = fcslib.Solver(fcslib.Mech(mech4), Treact=500.0)
    eq = fcslib.equilibrator(s)
    fields = [M.counterflow_field(mech4, 512, seed=sd, equilibrate=eq, n_rows=2) for sd in (31, 32, 33)]
    Xs, Ts = zip(*(learned.samples(s, f, dt) for f in fields[:2]))
    net, hist = learned.train(np.concatenate(Xs), np.concatenate(Ts), hidden=(256, 256, 256), epochs=40,
                              batch=256, lr=2e-3, seed=1)
    assert hist[-1] < 0.5 * hist[0]
    f = fields[2]

    def run(mlp=None):
        b = fcslib.Batch(s, f)
        if mlp is not None:
            b.set_learned(mlp, "verify")
        b.run(dt)
        return b.download(), b.learned_stats()
    plain, _ = run()
    lrn, (npred, nver, nacc) = run(net)
    react = f.T > 500.0
    acc = (lrn["status"] & 8) != 0
    assert npred == react.sum() and acc.sum() == nacc
    fb = react & ~acc
    for k in ("RR", "Tend", "deltaTChem", "nAttempts", "status"):
        a, b = lrn[k], plain[k]
        np.testing.assert_array_equal(a[..., fb], b[..., fb])
    rho0 = f.rho(mech4)
    Ya, Yb = f.Y + lrn["RR"] * dt / rho0, f.Y + plain["RR"] * dt / rho0
    d = np.abs(Ya - Yb)
    ok = ((d <= 1e-6) | (d <= 1e-4 * np.abs(Yb))).all(axis=0) & (np.abs(lrn["Tend"] - plain["Tend"]) <= 1e-3)
    assert ok[react].all()
    prod = f.cls == 1
    assert acc[prod].sum() >= 0.5 * prod.sum(), (acc[prod].sum(), prod.sum(), nver, nacc)
    assert lrn["nAttempts"][prod].sum() + (nver - nacc) <= 0.75 * plain["nAttempts"][prod].sum()


def test_large_batch_graph_autotune_lts_and_cold_cells(fcslib, mech80):
    """Workgroup form through the batch API: hipGraph replays and an
    autotuned launch shape give bit-identical results to eager launches
    (the matrix slices are regrown for larger grids), cold cells are skipped,
    and the local-time-step solve matches the oracle."""
    fld = M.uniform_field(mech80, 96, seed=44)
    fld.T[:8] = 350.0
    s = fcslib.Solver(fcslib.Mech(mech80), Treact=500.0)
    b = fcslib.Batch(s, fld)
    b.run(1e-6)
    ref = b.download()
    assert (ref["status"][:8] == 1).all() and (ref["RR"][:, :8] == 0).all()
    g = fcslib.Batch(s, fld)
    g.set_launch(0, 12)
    for _ in range(2):
        g.run(1e-6)
        r = g.download()
        for k in ("RR", "deltaTChem", "Tend", "status", "nAttempts"):
            np.testing.assert_array_equal(r[k], ref[k])
    grid, graph = b.autotune(1e-6)
    assert grid > 0 and graph in (0, 24)
    b.run(1e-6)
    r = b.download()
    for k in ("RR", "deltaTChem", "Tend", "status", "nAttempts"):
        np.testing.assert_array_equal(r[k], ref[k])
    dtf = 1e-6 * (1.0 + (np.arange(fld.n_cells) % 3))
    sub = fld.subset(np.arange(8, 24))
    gl = s.solve(sub, dtf[8:24])
    ol = Oracle(mech80).solve(sub, dtf[8:24], threads=1, Treact=500.0)
    assert gl["rc"] == 0 and ol["rc"] == 0
    Yg = sub.Y + gl["RR"] * dtf[8:24] / gl["rho0"]
    Yo = sub.Y + ol["RR"] * dtf[8:24] / gl["rho0"]
    d = np.abs(Yg - Yo)
    assert ((d <= 1e-6) | (d <= 1e-4 * np.abs(Yo))).all()
    assert np.abs(gl["Tend"] - ol["Tend"]).max() <= 1e-3


def test_reaction_tables_beyond_lds_take_the_hbm_matrix(fcslib):
    """70 species and 5,000 reactions: with the matrix in LDS the 80-, 96-
    and 112-row sizes would need more than a CU's 160 KiB, so the solver
    takes the 128-row size with the matrix in HBM (creation would otherwise
    fail with FCS_E_UNSUPPORTED); its solve matches the oracle."""
    mech = M.synthetic_mechanism(70, 5000, n_third_body=60, n_troe=20, n_plog=4, seed=7050)
    fld = M.uniform_field(mech, 8, seed=5)
    dt = 1e-7
    g = fcslib.Solver(fcslib.Mech(mech)).solve(fld, dt)
    o = Oracle(mech)
    r = o.solve(fld, dt, threads=8)
    assert g["rc"] == 0 and r["rc"] == 0
    assert_solve_parity(g, r, fld, dt, rerun=lambda f: o.solve(f, dt, threads=8))
