"""GPU parity on a real mechanism: GRI-3.0 (53 species, 325 reactions) from
the reference's counterflow tutorial and 1,000 cells of its converged
flame snapshot, through the C ABI, against the oracle."""
import numpy as np
import pytest

from oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fcslib():
    from fcs import lib
    lib.load()
    assert lib.device_count() >= 1
    return lib


def end_Y(field, r, dt):
    return field.Y + r["RR"] * dt / r["rho0"][None, :]


@pytest.mark.parametrize("solver", ["OptRodas34", "OptRosenbrock34", "OptSeulex"])
def test_flame_snapshot_solve(fcslib, gri30, solver):
    mech, field = gri30
    dt = 1e-6
    g = fcslib.Solver(fcslib.Mech(mech), solver=solver).solve(field, dt)
    o = Oracle(mech).solve(field, dt, threads=8, solver=solver)
    assert g["rc"] == 0 and o["rc"] == 0
    assert (g["status"] == o["status"]).all()
    dY = np.abs(end_Y(field, g, dt) - end_Y(field, o, dt))
    ok = (dY <= 1e-6) | (dY <= 1e-4 * np.abs(end_Y(field, o, dt)))
    assert ok.all(), dY.max()
    assert np.abs(g["Tend"] - o["Tend"]).max() <= 1e-3
    assert g["deltaTMin"] == pytest.approx(o["deltaTMin"], rel=0.2)


def test_flame_snapshot_kernels(fcslib, gri30):
    mech, field = gri30
    o = Oracle(mech)
    s = fcslib.Solver(fcslib.Mech(mech))
    idx = np.arange(0, field.n_cells, 16)
    YT = np.vstack([field.Y[:, idx], field.T[None, idx]])
    p = field.p[idx]
    d, _ = s.derivatives(p, YT)
    f, J, _ = s.jacobian(p, YT)
    for k in range(idx.size):
        dr, _ = o.derivatives(p[k], YT[:, k])
        scale = np.abs(dr[:-1]).max() + 1e-300
        assert np.abs(d[:-1, k] - dr[:-1]).max() <= 1e-11 * scale
        fr, Jr, _ = o.jacobian(p[k], YT[:, k], "exact")
        rs = np.abs(Jr).max(axis=1, keepdims=True) + 1e-300
        assert (np.abs(J[k] - Jr) / rs).max() <= 1e-9


def test_ignition_initial_state_steps(fcslib, gri30):
    """The 0-D ignition tutorial's CH4/air state (mole fractions CH4 0.5,
    O2 1, N2 3.76; 1000 K; 13.68 bar) integrated over ten 1e-5 s flow steps
    with the state carried between steps, GPU against oracle."""
    from fcs.mechanism import Field
    mech, _ = gri30
    X = np.zeros(mech.n_species)
    for n, v in (("CH4", 0.5), ("O2", 1.0), ("N2", 3.76)):
        X[mech.names.index(n)] = v
    Y = X * mech.W / (X * mech.W).sum()
    fg = Field(np.array([1000.0]), np.array([1.36789e6]), Y[:, None].copy(), np.array([1e-7]))
    fo = Field(fg.T.copy(), fg.p.copy(), fg.Y.copy(), fg.deltaTChem.copy())
    s = fcslib.Solver(fcslib.Mech(mech))
    o = Oracle(mech)
    for step in range(10):
        rg = s.solve(fg, 1e-5)
        ro = o.solve(fo, 1e-5)
        assert rg["rc"] == 0 and ro["rc"] == 0
        for f, r in ((fg, rg), (fo, ro)):
            f.Y = np.maximum(end_Y(f, r, 1e-5), 0.0)
            f.T = r["Tend"].copy()
            f.deltaTChem = r["deltaTChem"].copy()
        assert abs(fg.T[0] - fo.T[0]) <= 1e-3, step
        dY = np.abs(fg.Y - fo.Y)
        assert ((dY <= 1e-6) | (dY <= 1e-4 * fo.Y)).all(), step
