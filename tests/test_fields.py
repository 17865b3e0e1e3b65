"""Synthetic cell fields of the benchmark configurations: the 512 x 512
counterflow field's class fractions and its near-equilibrium product cells
(checked on a 32 x 32 field with the oracle as the equilibration step)."""
import numpy as np

import fcs.mechanism as M
from oracle import Oracle


def oracle_equilibrator(mech, deltaT=4e-3):
    o = Oracle(mech)

    def run(field):
        r = o.solve(field, deltaT)
        assert r["rc"] == 0
        Y = np.maximum(field.Y + r["RR"] * deltaT / r["rho0"][None, :], 0.0)
        return M.Field(r["Tend"].copy(), field.p.copy(), Y / Y.sum(axis=0), field.deltaTChem.copy())
    return run


def test_counterflow_field_classes_and_product_cells(mech1):
    # class fractions of the full 512 x 512 field (the equilibration step
    # does not change them): 45 % cold, 40 % products, 15 % flame front
    full = M.config_field(2, mech1, equilibrate=lambda f: f)
    assert full.n_cells == 262144
    frac = np.bincount(full.cls, minlength=3) / full.n_cells
    np.testing.assert_allclose(frac, [230 / 512, 206 / 512, 76 / 512], atol=1e-12)
    fld = M.config_field(2, mech1, n_cells=32 * 32, equilibrate=oracle_equilibrator(mech1))
    assert fld.n_cells == 1024
    cls = fld.cls
    assert (fld.T[cls == 0] < 500.0).all()
    assert (fld.T[cls == 2] >= 900.0).all() and (fld.T[cls == 2] <= 1800.0).all()
    np.testing.assert_allclose(fld.Y.sum(axis=0), 1.0, rtol=1e-12)
    # product cells are near equilibrium: a 1 us step takes only a few attempts,
    # flame-front cells take many
    o = Oracle(mech1)
    hot = np.nonzero(cls == 1)[0][:6]
    flame = np.nonzero(cls == 2)[0][:3]
    rh = o.solve(fld.subset(hot), 1e-6, Treact=500.0)
    rf = o.solve(fld.subset(flame), 1e-6, Treact=500.0)
    assert rh["rc"] == 0 and rf["rc"] == 0
    assert rh["nAttempts"].max() <= 4
    assert rf["nAttempts"].min() >= 40


def test_counterflow_field_needs_square_and_equilibrator(mech1):
    import pytest
    with pytest.raises(ValueError):
        M.config_field(2, mech1, n_cells=1000, equilibrate=lambda f: f)
    with pytest.raises(ValueError):
        M.config_field(2, mech1, n_cells=1024)
