"""Parity of the gfx950 library (through the C ABI) with the CPU oracle.

Bar: per-species |dY| <= 1e-6 absolute or 1e-4 relative and |dT| <= 1e-3 K
after solve(deltaT); kernel-level outputs at FP64 round-off scaled by the
row magnitude.
"""
import os

import numpy as np
import pytest

import fcs.mechanism as M
from oracle import Oracle, lu_solve as orc_lu_solve
from parity import assert_solve_parity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fcslib():
    from fcs import lib
    lib.load()
    assert lib.device_count() >= 1, "no GPU visible to the HIP runtime"
    return lib


def states(mech, cfg, n):
    fld = M.config_field(cfg, mech, n_cells=max(n, 8))
    YT = np.vstack([fld.Y[:, :n], fld.T[None, :n]])
    return fld.p[:n], YT


# ------------------------------------------------------------ kernel level
@pytest.mark.parametrize("cfg", [0, 1])
def test_derivatives(fcslib, cfg):
    mech = M.config_mechanism(cfg)
    o = Oracle(mech)
    s = fcslib.Solver(fcslib.Mech(mech))
    p, YT = states(mech, cfg, 64)
    YT[3, :8] = -1e-3      # exercises the in-place clip
    d, yc = s.derivatives(p, YT)
    for c in range(64):
        dref, yref = o.derivatives(p[c], YT[:, c])
        np.testing.assert_array_equal(yc[:, c], yref)
        scale = np.abs(dref[:-1]).max()
        assert np.abs(d[:-1, c] - dref[:-1]).max() <= 1e-11 * scale
        assert d[-1, c] == pytest.approx(dref[-1], rel=1e-10, abs=1e-11 * scale)


@pytest.mark.parametrize("cfg", [0, 1])
@pytest.mark.parametrize("jac", ["exact", "fast"])
def test_jacobian(fcslib, cfg, jac):
    mech = M.config_mechanism(cfg)
    o = Oracle(mech)
    s = fcslib.Solver(fcslib.Mech(mech), jacobian=jac)
    p, YT = states(mech, cfg, 32)
    f, J, _ = s.jacobian(p, YT)
    for c in range(32):
        fr, Jr, _ = o.jacobian(p[c], YT[:, c], jac)
        np.testing.assert_allclose(f[:, c], fr, rtol=1e-10, atol=1e-11 * np.abs(fr[:-1]).max())
        # the GPU uses the rank-1 form of the exact conversion: compare per row
        scale = np.abs(Jr).max(axis=1, keepdims=True) + 1e-300
        assert (np.abs(J[c] - Jr) / scale).max() <= 1e-9


@pytest.mark.parametrize("cfg", [0, 1])
@pytest.mark.parametrize("solver", ["OptRodas34", "OptRosenbrock34", "OptRodas23", "OptRosenbrock23"])
def test_single_attempt(fcslib, cfg, solver):
    """One Rodas34Solve / Rosenbrock34Solve / Rodas23Solve / Rosenbrock23Solve per state, from accepted-size
    steps (1e-14 s) to heavily rejected ones (1e-9 s): error norm and
    candidate state against the oracle.  Rejected steps amplify round-off,
    so the state bar is relative to the step's own increment."""
    mech = M.config_mechanism(cfg)
    o = Oracle(mech)
    s = fcslib.Solver(fcslib.Mech(mech), solver=solver)
    p, YT = states(mech, cfg, 24)
    YT[2, :4] = -1e-9      # start-state clip
    dx = np.repeat([1e-14, 1e-13, 1e-11, 1e-9], 6)
    err, YTn, yc = s.attempt(p, YT, dx)
    for c in range(24):
        er, ynr, yr = o.attempt(p[c], YT[:, c], dx[c], solver=solver)
        np.testing.assert_array_equal(yc[:, c], yr)
        tight = er < 1.0
        # the error norm is itself a cancellation product: relative bar 1e-4 with an absolute floor
        assert err[c] == pytest.approx(er, rel=1e-4, abs=1e-6), (c, err[c], er)
        inc = np.abs(ynr - yr).max()
        assert np.abs(YTn[:, c] - ynr).max() <= (1e-9 if tight else 1e-4) * inc + 1e-15, c


@pytest.mark.parametrize("n_species", [10, 53])
def test_lu_against_reference_golden(fcslib, n_species):
    """The device block-4 LU reproduces the reference LUsolver's solutions
    (tests/golden/lu_golden.npz, reference code compiled as-is)."""
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "lu_golden.npz"))
    mech = M.config_mechanism(0 if n_species == 10 else 1)
    n = mech.n_species + 1
    s = fcslib.Solver(fcslib.Mech(mech))
    A, b, x = [], [], []
    for i, sz in enumerate(g["sizes"]):
        if sz == n and i % 4 != 3:
            A.append(g["A"][g["moff"][i]:g["moff"][i + 1]].reshape(n, n))
            b.append(g["b"][g["voff"][i]:g["voff"][i + 1]])
            x.append(g["x"][g["voff"][i]:g["voff"][i + 1]])
    assert A
    xg = s.lu_solve(np.array(A), np.array(b))
    for k in range(len(A)):
        np.testing.assert_allclose(xg[k], x[k], rtol=0, atol=1e-10 * np.abs(x[k]).max())


def test_lu_pivot_quirk_on_device(fcslib, mech0):
    """Signed running maximum in the second column of a block: the device
    takes the same row as the reference (solution equals the oracle's)."""
    n = mech0.n_species + 1
    s = fcslib.Solver(fcslib.Mech(mech0))
    rng = np.random.default_rng(3)
    A = np.eye(n) + 0.01 * rng.standard_normal((n, n))
    A[1:4, 0] = 0.0
    A[1, 1], A[2, 1], A[3, 1] = 1.0, -5.0, 2.0
    b = rng.standard_normal(n)
    x = s.lu_solve(A[None], b[None])[0]
    xo, _, piv = orc_lu_solve(A, b)
    assert piv[1] == 3
    np.testing.assert_allclose(x, xo, rtol=1e-13, atol=1e-13)


# ------------------------------------------------------------ solve(deltaT)
@pytest.mark.parametrize("solver", ["OptRodas34", "OptRosenbrock34", "OptSeulex", "OptRodas23", "OptRosenbrock23"])
def test_solve_config0_all_cells(fcslib, mech0, solver):
    field = M.config_field(0, mech0)
    dt = M.CONFIG_DT[0]
    g = fcslib.Solver(fcslib.Mech(mech0), solver=solver).solve(field, dt)
    o = Oracle(mech0).solve(field, dt, threads=1, solver=solver)
    assert g["rc"] == 0 and o["rc"] == 0
    assert_solve_parity(g, o, field, dt)


def test_solve_config0_fast_jacobian(fcslib, mech0):
    field = M.config_field(0, mech0, n_cells=256)
    dt = M.CONFIG_DT[0]
    g = fcslib.Solver(fcslib.Mech(mech0), jacobian="fast").solve(field, dt)
    o = Oracle(mech0).solve(field, dt, threads=1, jacobian="fast")
    assert_solve_parity(g, o, field, dt)


@pytest.mark.parametrize("solver", ["OptRodas34", "OptSeulex", "OptRodas23", "OptRosenbrock23"])
def test_solve_config1_sample(fcslib, mech1, solver):
    field = M.config_field(1, mech1).subset(np.random.default_rng(1).choice(65536, 192, replace=False))
    dt = M.CONFIG_DT[1]
    g = fcslib.Solver(fcslib.Mech(mech1), solver=solver).solve(field, dt)
    orc = Oracle(mech1)
    o = orc.solve(field, dt, threads=1, solver=solver)
    assert g["rc"] == 0 and o["rc"] == 0
    # deltaTMin follows the smallest suggested step, which can jump between
    # discrete values with round-off (OptSeulex on this sample: cell 22's
    # 1.108e-8 s or cell 2's 1.347e-8 s, the oracle itself giving either
    # under 1e-14 relative perturbations of Y): the rerun accepts a value
    # the oracle reaches so
    assert_solve_parity(g, o, field, dt, rerun=lambda f2: orc.solve(f2, dt, threads=8, solver=solver))


def test_full_size_invariants_and_determinism(fcslib, mech1):
    """65,536 cells: every cell integrates, mass is conserved cell by cell,
    and two runs give bit-identical results."""
    field = M.config_field(1, mech1)
    s = fcslib.Solver(fcslib.Mech(mech1))
    b = fcslib.Batch(s, field)
    b.run(1e-6)
    r1 = b.download()
    b.run(1e-6)
    r2 = b.download()
    assert (r1["status"] == 0).all()
    for k in ("RR", "deltaTChem", "Tend", "nAttempts"):
        np.testing.assert_array_equal(r1[k], r2[k])
    # sum(Y) is conserved up to the clip of slightly negative mass fractions
    # at the end of the step (OptRodas34.C:114-117)
    rho0 = field.rho(mech1)
    assert (np.abs(r1["RR"].sum(axis=0)) * 1e-6 / rho0).max() <= 1e-8


# ------------------------------------------------------------ edge cases
def test_cold_cells_cap_and_batch_api(fcslib, mech0):
    field = M.config_field(0, mech0, n_cells=128)
    field.T[::3] = 450.0
    dt = 1e-5
    s = fcslib.Solver(fcslib.Mech(mech0), Treact=500.0, deltaTChemMax=2e-6)
    g = s.solve(field, dt)
    o = Oracle(mech0).solve(field, dt, threads=1, Treact=500.0, deltaTChemMax=2e-6)
    assert_solve_parity(g, o, field, dt)
    cold = field.T <= 500.0
    assert (g["status"][cold] == 1).all() and (g["RR"][:, cold] == 0).all()
    np.testing.assert_array_equal(g["deltaTChem"][cold], field.deltaTChem[cold])
    assert (g["deltaTChem"][~cold] <= 2e-6).all()
    b = fcslib.Batch(s, field)
    b.run(dt)
    r = b.download()
    for k in ("RR", "deltaTChem", "Tend", "status"):
        np.testing.assert_array_equal(r[k], g[k])
    assert r["deltaTMin"] == g["deltaTMin"]


def test_maxsteps_and_nonfinite_are_reported(fcslib, mech0):
    from fcs.abi import FCS_E_CELLS
    field = M.config_field(0, mech0, n_cells=8)
    s = fcslib.Solver(fcslib.Mech(mech0), maxSteps=3)
    g = s.solve(field, 1e-5)
    assert g["rc"] == FCS_E_CELLS and (g["status"] & 2).all()
    # T = inf reaches the rates through the unclamped density: the attempt's
    # error norm is NaN and the cell is reported non-finite, like the oracle
    field2 = M.config_field(0, mech0, n_cells=8)
    field2.T[2] = np.inf
    g2 = fcslib.Solver(fcslib.Mech(mech0)).solve(field2, 1e-5)
    o2 = Oracle(mech0).solve(field2, 1e-5)
    assert g2["rc"] == FCS_E_CELLS and g2["status"][2] & 4
    assert (g2["status"] == o2["status"]).all()
    assert g2["nAttempts"][2] == o2["nAttempts"][2]
    # a NaN mass fraction is clipped to 0 by y0 = max(0, Y) (OptRodas34.C:72-81):
    # the cell integrates, and only its RR (which subtracts Y0) is NaN
    field3 = M.config_field(0, mech0, n_cells=8)
    field3.Y[0, 2] = np.nan
    g3 = fcslib.Solver(fcslib.Mech(mech0)).solve(field3, 1e-5)
    assert g3["rc"] == 0 and (g3["status"] == 0).all()
    np.testing.assert_array_equal(np.isnan(g3["RR"]).any(axis=0), np.arange(8) == 2)


def test_seulex_batch_and_fast_jacobian(fcslib, mech0):
    """OptSeulex through the device-resident batch (the per-cell Jacobian is
    kept in HBM across steps) with jacobian fast, against the oracle and
    against fcs_solve; the single-attempt entry has no Seulex form."""
    from fcs.abi import FCS_E_UNSUPPORTED
    from fcs.lib import FcsError
    field = M.config_field(0, mech0, n_cells=256)
    dt = M.CONFIG_DT[0]
    s = fcslib.Solver(fcslib.Mech(mech0), solver="OptSeulex", jacobian="fast")
    g = s.solve(field, dt)
    o = Oracle(mech0).solve(field, dt, threads=1, solver="OptSeulex", jacobian="fast")
    assert g["rc"] == 0 and o["rc"] == 0
    assert_solve_parity(g, o, field, dt)
    # column counts follow the same decisions except where round-off tips one
    assert np.mean(g["nAttempts"] == o["nAttempts"]) >= 0.9
    b = fcslib.Batch(s, field)
    b.run(dt)
    r = b.download()
    for k in ("RR", "deltaTChem", "Tend", "status", "nAttempts"):
        np.testing.assert_array_equal(r[k], g[k])
    with pytest.raises(FcsError) as e:
        s.attempt(field.p[:4], np.vstack([field.Y[:, :4], field.T[None, :4]]), np.full(4, 1e-9))
    assert e.value.code == FCS_E_UNSUPPORTED


def test_empty_batch(fcslib, mech0):
    s = fcslib.Solver(fcslib.Mech(mech0))
    field = M.config_field(0, mech0, n_cells=8).subset(np.arange(0))
    g = s.solve(field, 1e-5)
    assert g["rc"] == 0 and g["deltaTMin"] == 2e-5


def test_solve_lts_per_cell_deltaT(fcslib, mech0):
    """Local-time-step solve(scalarField deltaT): per-cell flow steps, RR
    formed with the cell's own deltaT, deltaTMin without the 2*deltaT cap."""
    field = M.config_field(0, mech0, n_cells=256)
    dt = np.random.default_rng(5).uniform(2e-6, 2e-5, 256)
    g = fcslib.Solver(fcslib.Mech(mech0)).solve(field, dt)
    o = Oracle(mech0).solve(field, dt, threads=1)
    assert g["rc"] == 0 and o["rc"] == 0
    assert_solve_parity(g, o, field, dt)


def test_graph_replays_match_eager_launches(fcslib, mech1):
    """Step launches captured in a hipGraph: every replay (with eager work
    and downloads in between) gives bit-identical results to eager
    launches; a new deltaT re-captures."""
    field = M.config_field(1, mech1, n_cells=2048)
    s = fcslib.Solver(fcslib.Mech(mech1))
    ref = {}
    for dt in (1e-6, 5e-7):
        b = fcslib.Batch(s, field)
        b.run(dt)
        ref[dt] = b.download()
    g = fcslib.Batch(s, field)
    g.set_launch(0, 12)
    for dt in (1e-6, 1e-6, 5e-7, 1e-6):
        g.run(dt)
        r = g.download()
        for k in ("RR", "deltaTChem", "Tend", "status", "nAttempts"):
            np.testing.assert_array_equal(r[k], ref[dt][k])
        assert r["deltaTMin"] == ref[dt]["deltaTMin"]


def test_autotuned_launch_shape_keeps_results(fcslib, mech0):
    field = M.config_field(0, mech0, n_cells=512)
    s = fcslib.Solver(fcslib.Mech(mech0))
    b = fcslib.Batch(s, field)
    b.run(1e-5)
    r0 = b.download()
    grid, graph = b.autotune(1e-5)
    assert grid > 0 and graph in (0, 24)
    b.run(1e-5)
    r1 = b.download()
    for k in ("RR", "deltaTChem", "Tend", "status", "nAttempts"):
        np.testing.assert_array_equal(r1[k], r0[k])


# ------------------------------------------------- SRI and nonEquilibrium
def sri_noneq_mech():
    """12 species, 48 reactions: 4 third-body, 2 Troe, 4 SRI falloff, 6
    nonEquilibrium (explicit reverse Arrhenius) reactions."""
    return M.synthetic_mechanism(12, 48, n_third_body=4, n_troe=2, n_sri=4, n_noneq=6, seed=4242)


def test_sri_noneq_kernels_and_solve(fcslib):
    mech = sri_noneq_mech()
    assert sum(r.rtype == M.RXN_SRI for r in mech.reactions) == 4
    assert sum(bool(r.rev_arr) for r in mech.reactions) == 6
    o = Oracle(mech)
    fld = M.uniform_field(mech, 128, seed=99)
    YT = np.vstack([fld.Y, fld.T[None, :]])
    s = fcslib.Solver(fcslib.Mech(mech))
    d, _ = s.derivatives(fld.p, YT)
    f, J, _ = s.jacobian(fld.p, YT)
    for c in range(0, 128, 4):
        dref, _ = o.derivatives(fld.p[c], YT[:, c])
        scale = np.abs(dref[:-1]).max()
        assert np.abs(d[:-1, c] - dref[:-1]).max() <= 1e-11 * scale
        assert d[-1, c] == pytest.approx(dref[-1], rel=1e-10, abs=1e-11 * scale)
        fr, Jr, _ = o.jacobian(fld.p[c], YT[:, c], "exact")
        np.testing.assert_allclose(f[:, c], fr, rtol=1e-10, atol=1e-11 * np.abs(fr[:-1]).max())
        rs = np.abs(Jr).max(axis=1, keepdims=True) + 1e-300
        assert (np.abs(J[c] - Jr) / rs).max() <= 1e-9
    dt = 1e-5
    g = s.solve(fld, dt)
    r = Oracle(mech).solve(fld, dt, threads=1)
    assert g["rc"] == 0 and r["rc"] == 0
    assert_solve_parity(g, r, fld, dt)


# ------------------------------------------------ calculate / tc / Qdot
@pytest.mark.parametrize("which", ["cfg0", "cfg1", "sri_noneq"])
def test_calculate_tc_qdot(fcslib, which):
    """fcs_calculate, fcs_tc and fcs_qdot against the oracle; the states
    include a negative mass fraction (no clipping in these paths) and a cold
    cell below the fit range (clamped temperature)."""
    if which == "sri_noneq":
        mech = sri_noneq_mech()
        fld = M.uniform_field(mech, 96, seed=5)
    else:
        cfg = int(which[-1])
        mech = M.config_mechanism(cfg)
        fld = M.config_field(cfg, mech, n_cells=96)
    Y = fld.Y.copy()
    Y[1, 3] = -1e-4
    T = fld.T.copy()
    T[5] = 150.0
    rho = fld.p / (M.RU * T * (np.abs(Y) / mech.W[:, None]).sum(axis=0))
    s = fcslib.Solver(fcslib.Mech(mech))
    o = Oracle(mech)
    RR = s.calculate(rho, T, fld.p, Y)
    tc = s.tc(rho, T, fld.p, Y)
    q = s.qdot(RR)
    for c in range(96):
        rr = o.calculate(rho[c], T[c], fld.p[c], Y[:, c])
        np.testing.assert_allclose(RR[:, c], rr, rtol=1e-10, atol=1e-11 * np.abs(rr).max())
        assert tc[c] == pytest.approx(o.tc(rho[c], T[c], fld.p[c], Y[:, c]), rel=1e-10)
        assert q[c] == pytest.approx(o.qdot(RR[:, c]), rel=1e-12, abs=1e-12 * np.abs(o.hf() * RR[:, c]).sum())
