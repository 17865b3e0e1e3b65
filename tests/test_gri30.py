"""The OpenFOAM dictionary loader and the oracle on a real mechanism:
GRI-3.0 (53 species, 325 reactions: Arrhenius, third-body, Lindemann and
Troe falloff, irreversible) from the reference's counterflow tutorial, with
states from its converged flame snapshot (tests/golden/gri30_counterflow.npz)."""
import os

import numpy as np
import pytest

from fcs import foam
from oracle import Oracle

REF_CASE = "/root/reference/tutorial/CounterFlowFlame2D/CH4_GRI30"


def test_dictionary_grammar():
    d = foam.parse('''
        FoamFile { format ascii; }
        // comment
        species 3 ( H2 CH2(S) O2 );   /* block
        comment */
        r0 { type reversibleThirdBodyArrhenius; reaction "2O = O2"; A 1.2e+11; beta -1; Ta 0;
             coeffs 2 ( (H2 2.4) (CH2(S) 1) ); }
        F { alpha 0.5; list ( 1 2 3 ); }
    ''')
    assert "FoamFile" not in d
    assert d["species"] == ["H2", "CH2(S)", "O2"]
    assert d["r0"]["reaction"] == "2O = O2" and d["r0"]["A"] == 1.2e11 and d["r0"]["beta"] == -1.0
    assert d["r0"]["coeffs"] == [["H2", 2.4], ["CH2(S)", 1.0]]
    assert d["F"] == {"alpha": 0.5, "list": [1.0, 2.0, 3.0]}


def test_reaction_sides_expand_integer_coefficients():
    idx = {"O": 0, "O2": 1, "CO": 2, "OH": 3}
    assert foam._side(" 2O ", idx) == [0, 0]
    assert foam._side("OH + 2CO", idx) == [3, 2, 2]
    with pytest.raises(ValueError):
        foam._side("X + O", idx)
    assert foam._side("2O + 2OH", idx) == [0, 0, 3, 3]   # long sides: checked per reaction form
    with pytest.raises(ValueError):
        foam._side("9O + 8OH", idx)


@pytest.mark.skipif(not os.path.isdir(REF_CASE), reason="reference tutorial not present")
def test_loader_reproduces_the_fixture(gri30):
    mech, _ = gri30
    m = foam.load_mechanism(os.path.join(REF_CASE, "constant", "reactions"),
                            os.path.join(REF_CASE, "constant", "speciesThermo"))
    a, b = m.arrays(), mech.arrays()
    for k in a:
        np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]), err_msg=k)


def test_gri30_composition(gri30):
    mech, field = gri30
    assert mech.n_species == 53 and mech.n_reactions == 325
    t = np.bincount([r.rtype for r in mech.reactions], minlength=5)
    assert list(t) == [284, 12, 3, 26, 0]
    assert sum(not r.reversible for r in mech.reactions) == 16
    assert field.n_cells == 1000 and 290 < field.T.min() and field.T.max() < 2100
    np.testing.assert_allclose(field.Y.sum(axis=0), 1.0, atol=2e-6)


def test_oracle_on_flame_states(gri30):
    mech, field = gri30
    o = Oracle(mech)
    rng = np.random.default_rng(0)
    for c in rng.choice(field.n_cells, 6, replace=False):
        YT = np.concatenate([field.Y[:, c], [field.T[c]]])
        d, _ = o.derivatives(field.p[c], YT)
        # mass conservation of the species rates, to the precision of the
        # tabulated molecular weights (7 significant digits: sum_i nu_ri W_i
        # is not exactly zero for every reaction)
        assert abs(d[:-1].sum()) <= 1e-6 * np.abs(d[:-1]).sum()
        # Jacobian against central finite differences of the RHS (states
        # lifted off the clip at Y = 0), per row relative to the row scale
        y = np.maximum(YT, 1e-3)
        y[:-1] /= y[:-1].sum()
        f, J, _ = o.jacobian(field.p[c], y, "exact")
        Jfd = np.zeros_like(J)
        for j in range(y.size):
            h = 1e-6 * max(abs(y[j]), 1e-4) if j < y.size - 1 else 1e-3
            yp, ym = y.copy(), y.copy()
            yp[j] += h
            ym[j] -= h
            Jfd[:, j] = (o.derivatives(field.p[c], yp)[0] - o.derivatives(field.p[c], ym)[0]) / (2 * h)
        scale = np.abs(Jfd).max(axis=1, keepdims=True) + 1e-300
        assert (np.abs(J - Jfd) / scale).max() < 2e-5, c
    r = o.solve(field, 1e-6, threads=8)
    assert r["rc"] == 0 and (r["status"] == 0).all()
