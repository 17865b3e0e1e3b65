"""The oracle's local-time-step solve (FastChemistryModel_localEulerSolve.H)
against its serial solve: same integration, RR formed as (Y - Y0)*rho0/dt
instead of (Y*rho0 - Y0*rho0)/dt, deltaTMin without the 2*deltaT cap."""
import numpy as np

import fcs.mechanism as M
from oracle import Oracle


def test_lts_with_a_uniform_field_matches_serial(mech0):
    field = M.config_field(0, mech0, n_cells=24)
    dt = M.CONFIG_DT[0]
    o = Oracle(mech0)
    s = o.solve(field, dt)
    l = o.solve(field, np.full(field.n_cells, dt))
    assert s["rc"] == 0 and l["rc"] == 0
    np.testing.assert_array_equal(s["deltaTChem"], l["deltaTChem"])
    np.testing.assert_array_equal(s["Tend"], l["Tend"])
    np.testing.assert_allclose(l["RR"], s["RR"], rtol=1e-12, atol=1e-12 * np.abs(s["RR"]).max())
    # serial: min(deltaTMin, 2 deltaT); LTS: the plain minimum
    raw = l["deltaTMin"]
    assert raw == s["deltaTChem"].min() or raw <= s["deltaTChem"].min()
    assert s["deltaTMin"] == min(raw, 2 * dt)


def test_lts_cells_use_their_own_deltaT(mech0):
    field = M.config_field(0, mech0, n_cells=8)
    o = Oracle(mech0)
    dts = np.array([1e-6, 2e-6, 4e-6, 8e-6, 1e-6, 2e-6, 4e-6, 8e-6])
    l = o.solve(field, dts)
    for c in range(8):
        one = o.solve(field.subset([c]), float(dts[c]))
        Y1 = field.Y[:, c] + one["RR"][:, 0] * dts[c] / one["rho0"][0]
        Yl = field.Y[:, c] + l["RR"][:, c] * dts[c] / l["rho0"][c]
        np.testing.assert_allclose(Yl, Y1, rtol=1e-12, atol=1e-15)
        assert l["Tend"][c] == one["Tend"][0]
    # no cell integrated: deltaTMin is OpenFOAM's great
    cold = field.subset(np.arange(2))
    cold.T[:] = 200.0
    r = o.solve(cold, np.full(2, 1e-6), Treact=500.0)
    assert r["deltaTMin"] == 1e15
