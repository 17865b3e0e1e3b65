"""Pins the oracle's steppers: order of accuracy of the Rosenbrock tableaus
(textbook autonomous form), accuracy of whole integrations against an
independent scipy Radau solution, controller step ratios, and the
solve(deltaT) bookkeeping contract."""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

import fcs.mechanism as M
from oracle import Oracle, lib as orc_lib

from test_oracle_thermo_rates import R, tiny_mech


def smooth_mech():
    """S0 + S1 -> S2 -> S3 (mass-balanced, S3 an isomer of S2) with time
    scales around 1e-4 s at 1200 K."""
    m = tiny_mech([R(M.RXN_ARRHENIUS, [0, 1], [2], A=3.7e7, beta=0.0, Ta=3000.0),
                   R(M.RXN_ARRHENIUS, [2], [3], A=5.3e4, beta=0.0, Ta=2000.0)], seed=5)
    m.W = np.array([10.0, 20.0, 30.0, 30.0])
    return m


def radau(o, p, y0, t_end, rtol=1e-12, atol=1e-16):
    f = lambda t, y: o.derivatives(p, y)[0]
    jac = lambda t, y: o.jacobian(p, y)[1]
    s = solve_ivp(f, (0.0, t_end), y0, method="Radau", rtol=rtol, atol=atol, jac=jac)
    assert s.success
    return s.y[:, -1]


def one_step(o, p, y0, h, solver):
    r = o.integrate_cell(p, y0, h, h, solver=solver, absTol=1e10, relTol=1e10, log_cap=4)
    assert r["rc"] == 0 and r["nAttempts"] == 1
    return r["y"], r["log"][0][2]


@pytest.mark.parametrize("solver,order", [("OptRodas34", 4), ("OptRosenbrock34", 4),
                                          ("OptRodas23", 3), ("OptRosenbrock23", 3)])
def test_rosenbrock_tableau_order(solver, order):
    """Textbook form (df/dt = 0): local error of the propagated solution is
    O(h^(order+1)) and the embedded error estimate is O(h^order)."""
    m = smooth_mech()
    o = Oracle(m)
    y0 = np.array([0.3, 0.3, 0.2, 0.2, 1200.0])
    p = 1e5
    orc_lib().orc_set_textbook_rosenbrock(1)
    try:
        errs, ests = [], []
        hs = [1e-5, 5e-6, 2.5e-6]
        for h in hs:
            y1, est = one_step(o, p, y0, h, solver)
            ref = radau(o, p, y0, h)
            errs.append(np.abs(y1[:-1] - ref[:-1]).max())
            ests.append(est * 1e10)  # norm = max|err| / (absTol + relTol max|y|)
        r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
        lo, hi = 2 ** (order + 0.5), 2 ** (order + 1.6)
        assert lo < r1 < hi and lo < r2 < hi, (errs, r1, r2)
        e1, e2 = ests[0] / ests[1], ests[1] / ests[2]
        assert lo / 2 < e1 < hi / 2 and lo / 2 < e2 < hi / 2, (ests, e1, e2)
    finally:
        orc_lib().orc_set_textbook_rosenbrock(0)


# Seulex keeps its order near 2-3 on this mechanism, and the third-order
# Rosenbrock methods are third order: at relTol 1e-10 they need more than
# maxSteps = 10000 steps, so they are checked at 1e-8.
@pytest.mark.parametrize("solver,atol,rtol,tolY", [("OptRodas34", 1e-14, 1e-10, 1e-7),
                                                  ("OptRosenbrock34", 1e-14, 1e-10, 1e-7),
                                                  ("OptRodas23", 1e-12, 1e-8, 1e-6),
                                                  ("OptRosenbrock23", 1e-12, 1e-8, 1e-6),
                                                  ("OptSeulex", 1e-12, 1e-8, 1e-6)])
def test_tight_tolerance_matches_radau(solver, atol, rtol, tolY):
    """Tight-tolerance integrations agree with an independent Radau solution."""
    m = M.config_mechanism(0)
    o = Oracle(m)
    fld = M.config_field(0, m, n_cells=4)
    for c in range(2):
        y0 = np.concatenate([fld.Y[:, c], [fld.T[c]]])
        r = o.integrate_cell(fld.p[c], y0, 1e-5, 1e-7, solver=solver, absTol=atol, relTol=rtol)
        assert r["rc"] == 0
        ref = radau(o, fld.p[c], y0, 1e-5)
        assert np.abs(r["y"][:-1] - np.maximum(ref[:-1], 0)).max() < tolY
        assert abs(r["y"][-1] - ref[-1]) < 100 * tolY * ref[-1]


@pytest.mark.parametrize("solver", ["OptRodas34", "OptRosenbrock34", "OptSeulex"])
def test_default_tolerance_error_is_bounded(solver):
    """absTol 1e-8 / relTol 1e-4: end state within 1e-4 relative of Radau."""
    m = M.config_mechanism(0)
    o = Oracle(m)
    fld = M.config_field(0, m, n_cells=4)
    for c in range(2):
        y0 = np.concatenate([fld.Y[:, c], [fld.T[c]]])
        r = o.integrate_cell(fld.p[c], y0, 1e-5, 1e-7, solver=solver)
        ref = radau(o, fld.p[c], y0, 1e-5)
        assert np.abs(r["y"][:-1] - np.maximum(ref[:-1], 0)).max() < 2e-4
        assert abs(r["y"][-1] - ref[-1]) < 1e-4 * ref[-1]


@pytest.mark.parametrize("solver", ["OptRodas34", "OptRosenbrock34"])
def test_controller_step_ratios(solver):
    """Rejections shrink the step by max(0.9 Err^-1/4, 0.2); accepted steps
    grow it by at most 10x (src/ODE/OptRodas34/OptRodas34.H:105-112)."""
    m = M.config_mechanism(0)
    o = Oracle(m)
    fld = M.config_field(0, m, n_cells=4)
    y0 = np.concatenate([fld.Y[:, 0], [fld.T[0]]])
    r = o.integrate_cell(fld.p[0], y0, 1e-5, 1e-7, solver=solver, log_cap=5000)
    log = r["log"]
    assert any(not a for _, _, _, a, _ in log), "expected at least one rejection"
    for (x0, dx0, e0, a0, _), (x1, dx1, e1, a1, _) in zip(log, log[1:]):
        if not a0:
            assert x1 == x0
            assert dx1 == pytest.approx(dx0 * max(0.9 * e0 ** -0.25, 0.2), rel=1e-12)
            assert e0 > 1
        else:
            assert e0 <= 1
            assert x1 == pytest.approx(x0 + dx0, rel=1e-14)
            grow = dx1 / dx0
            if x1 + dx1 < 1e-5 * (1 - 1e-12):     # not truncated at the end
                assert 0.2 - 1e-12 <= grow <= 10 + 1e-12
    assert sum(d for _, d, _, a, _ in log if a) == pytest.approx(1e-5, rel=1e-12)


def test_solve_contract(mech0):
    """RR = (Y - Y0) rho0 / dt; cold cells skipped with RR = 0 and deltaTChem
    untouched; deltaTMin = min(min deltaTChem, 2 dt); deltaTChem capped."""
    o = Oracle(mech0)
    fld = M.config_field(0, mech0, n_cells=16)
    fld.T[:4] = 400.0
    dt = 1e-5
    res = o.solve(fld, dt, Treact=500.0, deltaTChemMax=3e-6)
    assert (res["status"][:4] == 1).all()
    np.testing.assert_array_equal(res["RR"][:, :4], 0.0)
    np.testing.assert_array_equal(res["deltaTChem"][:4], fld.deltaTChem[:4])
    assert (res["deltaTChem"][4:] <= 3e-6).all()
    for c in range(4, 16):
        y0 = np.concatenate([fld.Y[:, c], [fld.T[c]]])
        r = o.integrate_cell(fld.p[c], y0, dt, fld.deltaTChem[c])
        np.testing.assert_allclose(res["RR"][:, c], (r["y"][:-1] * res["rho0"][c] - fld.Y[:, c] * res["rho0"][c]) / dt,
                                   rtol=1e-14, atol=1e-300)
        assert res["Tend"][c] == r["y"][-1]
    assert res["deltaTMin"] <= 2 * dt
    # mass is conserved cell by cell
    assert np.abs(res["RR"][:, 4:].sum(axis=0)).max() < 1e-8 * np.abs(res["RR"]).max()
