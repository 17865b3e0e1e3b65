"""The C++ chemistry-model adapter (include/fcs_chemistry_model.hpp) through
its host harness (csrc/fcs_harness.cpp), which stands in for OpenFOAM: option
reading from a chemistryProperties dictionary (CPU), and on the GPU a time
loop of solve(deltaT) calls with the species advanced by the returned rates,
checked against the same steps driven from Python and against the oracle."""
import os

import numpy as np
import pytest

import fcs.mechanism as M
from fcs import foam, harness

CHEMISTRY_PROPERTIES = """
FoamFile { version 2; format ascii; class dictionary; object chemistryProperties; }
chemistryType
{
    solver          OptRodas34;
    method          FastChemistryModel;
}
chemistry       on;
Treact          500;
jacobian        fast;
initialChemicalTimeStep 1e-07;
maxChemicalTimeStep 1e-4;
OptRodas34Coeffs
{
    absTol          1e-8;
    relTol          1e-4;
    maxSteps        20000;
}
OptSeulexCoeffs
{
    absTol          1e-9;
    relTol          1e-5;
}
"""


def case(tmp_path, text, mech=None, field=None):
    mech = mech or M.config_mechanism(0)
    field = field or M.config_field(0, mech, n_cells=8)
    d = str(tmp_path / "case")
    harness.write_case(d, mech, field, harness.flatten(foam.parse(text)))
    return d


def test_adapter_reads_chemistry_properties(tmp_path):
    o = harness.options(case(tmp_path, CHEMISTRY_PROPERTIES))
    assert o["solver"] == "OptRodas34" and o["solver_id"] == 0
    assert o["jacobian"] == 1 and o["Treact"] == 500.0
    assert o["absTol"] == 1e-8 and o["relTol"] == 1e-4 and o["maxSteps"] == 20000
    assert o["initialChemicalTimeStep"] == 1e-7 and o["deltaTChemMax"] == 1e-4 and o["chemistry"] == 1


def test_adapter_defaults_and_other_solvers(tmp_path):
    text = CHEMISTRY_PROPERTIES.replace("solver          OptRodas34;", "solver          OptSeulex;")
    text = text.replace("jacobian        fast;", "").replace("maxChemicalTimeStep 1e-4;", "")
    text = text.replace("Treact          500;", "").replace("initialChemicalTimeStep 1e-07;", "")
    o = harness.options(case(tmp_path, text))
    assert o["solver"] == "OptSeulex" and o["solver_id"] == 2 and o["jacobian"] == 0
    assert o["absTol"] == 1e-9 and o["relTol"] == 1e-5 and o["maxSteps"] == 10000
    assert o["Treact"] == 0.0 and o["deltaTChemMax"] == 1e15 and o["initialChemicalTimeStep"] == 1e15


@pytest.mark.parametrize("edit,message", [
    (("solver          OptRodas34;", "solver          OptRodas99;"), "Unknown chemistry solver OptRodas99"),
    (("jacobian        fast;", "jacobian        approximate;"), "jacobian must be exact or fast"),
    (("    relTol          1e-4;\n    maxSteps", "    maxSteps"), "mandatory in OptRodas34Coeffs"),
    (("method          FastChemistryModel;", "method          standard;"), "must be FastChemistryModel"),
])
def test_adapter_rejects_bad_options(tmp_path, edit, message):
    text = CHEMISTRY_PROPERTIES.replace(*edit)
    with pytest.raises(RuntimeError, match=message):
        harness.options(case(tmp_path, text))


@pytest.mark.gpu
@pytest.mark.parametrize("lts", [False, True])
def test_adapter_time_loop_matches_library_and_oracle(tmp_path, lts):
    """Three flow steps through the adapter: the rates, deltaTChem and
    deltaTMin equal the same steps driven through the Python binding (same
    library, bit for bit), the first step agrees with the oracle, and
    calculate / tc / Qdot agree with the library at the end state."""
    from fcs import lib
    from oracle import Oracle
    mech = M.config_mechanism(0)
    field = M.config_field(0, mech, n_cells=64)
    field.T[:4] = 400.0   # below Treact: skipped cells
    text = CHEMISTRY_PROPERTIES.replace("jacobian        fast;", "jacobian        exact;")
    d = case(tmp_path, text, mech, field)
    dt, n_steps = 1e-5, 3
    out = harness.run(d, dt, n_steps, mech.n_species, field.n_cells, lts=lts)

    s = lib.Solver(lib.Mech(mech), jacobian="exact", absTol=1e-8, relTol=1e-4, maxSteps=20000, Treact=500.0,
                   deltaTChemMax=1e-4)
    rho = field.rho(mech)
    dtf = dt * (1.0 + (np.arange(field.n_cells) % 3)) if lts else dt
    f = M.Field(field.T.copy(), field.p.copy(), field.Y.copy(), np.full(field.n_cells, 1e-7))
    dmins = []
    for step in range(n_steps):
        r = s.solve(f, dtf, rho=rho)
        if step == 0:
            o = Oracle(mech).solve(f, dtf, threads=1, rho=rho, jacobian="exact", Treact=500.0,
                                   deltaTChemMax=1e-4, maxSteps=20000)
            ok = (np.abs(r["RR"] - o["RR"]) * np.asarray(dtf) / rho <= 1e-6) | \
                 (np.abs(r["RR"] - o["RR"]) <= 1e-4 * np.abs(o["RR"]))
            assert ok.all()
        dmins.append(r["deltaTMin"])
        f = M.Field(f.T, f.p, f.Y + np.asarray(dtf) * r["RR"] / rho, r["deltaTChem"])
    np.testing.assert_array_equal(out["deltaTMin"], dmins)
    np.testing.assert_array_equal(out["RR"], r["RR"])
    np.testing.assert_array_equal(out["deltaTChem"], f.deltaTChem)
    np.testing.assert_array_equal(out["Y"], f.Y)
    assert (out["RR"][:, :4] == 0).all()
    np.testing.assert_array_equal(out["RR_calculate"], s.calculate(rho, f.T, f.p, f.Y))
    np.testing.assert_array_equal(out["tc"], s.tc(rho, f.T, f.p, f.Y))
    np.testing.assert_array_equal(out["Qdot"], s.qdot(r["RR"]))


@pytest.mark.gpu
def test_adapter_reads_long_product_sides(tmp_path, tutorial_case):
    """A tutorial mechanism with 5-7-slot product sides (slot stride 7 in
    mech.bin) through the harness: the rates equal the library's."""
    from fcs import lib
    mech, field = tutorial_case("Kerosene_Zettervall2020")
    assert mech.arrays()["slot_stride"] == 7
    field = field.subset(np.arange(16))
    text = CHEMISTRY_PROPERTIES.replace("jacobian        fast;", "jacobian        exact;")
    d = case(tmp_path, text, mech, field)
    out = harness.run(d, 1e-5, 1, mech.n_species, field.n_cells)
    s = lib.Solver(lib.Mech(mech), jacobian="exact", absTol=1e-8, relTol=1e-4, maxSteps=20000, Treact=500.0,
                   deltaTChemMax=1e-4)
    f = M.Field(field.T.copy(), field.p.copy(), field.Y.copy(), np.full(field.n_cells, 1e-7))
    r = s.solve(f, 1e-5, rho=field.rho(mech))
    np.testing.assert_array_equal(out["RR"], r["RR"])
    np.testing.assert_array_equal(out["deltaTMin"], [r["deltaTMin"]])
