"""Mechanisms of the reference's tutorials (tests/golden/tutorial_mechs.npz,
written by tests/golden/make_tutorial_mechs.py): what the loader reads from
the files, and the oracle on the forms the synthetic mechanisms lack --
ArrheniusPLOG rows, reactions with 4 to 7 product slots, irreversible ones and
a Kc-reversible one (the reference's general path RFG / JFG, ReactionFunctionI.H:648-720,
JacobianFunctionI.H:2965-3060, ReactionRate.C:565-569) and the
nonEquilibrium rows of real mechanisms.

Pins: numbers typed from the tutorial files; the closed form of a single
reaction's rates; mass conservation (the tutorials' molecular weights are
element sums, so every reaction balances to round-off and any dropped or
extra slot shows); central finite differences of the RHS for the Jacobian.
"""
import copy

import numpy as np
import pytest

import fcs.mechanism as M
from fcs import foam
from oracle import Oracle

from conftest import TUTORIAL_CASES

# species, reactions, {form: count}, nonEquilibrium rows, most product slots
EXPECTED = {
    "NH3_Zhang_counterflow": (37, 261, {0: 232, 1: 15, 2: 1, 3: 12, 4: 1}, 0, 3),
    "NC12H26_Yao_counterflow": (53, 268, {0: 237, 1: 5, 2: 1, 3: 25}, 0, 7),
    "H2_Connaire2004": (10, 21, {0: 15, 1: 4, 3: 2}, 19, 2),
    "Kerosene_Zettervall2020": (31, 77, {0: 65, 1: 10, 2: 2}, 0, 7),
    "NC12H26_Gao2014": (31, 193, {0: 170, 1: 5, 2: 1, 3: 17}, 0, 4),
    "DME_LLNL": (79, 351, {0: 314, 1: 18, 2: 5, 3: 14}, 332, 3),
    "NH3CH4_Konnov": (129, 1231, {0: 1140, 1: 54, 2: 13, 3: 22, 5: 2}, 0, 3),
    "NC7H16_Yoo2011": (188, 939, {0: 905, 1: 19, 2: 5, 3: 10}, 924, 3),
    "NC12H26_Narayanaswamy2014": (255, 1509, {0: 1465, 1: 10, 3: 34}, 3, 4),
    "C2H4_Sandiego": (59, 271, {0: 233, 1: 15, 3: 23}, 0, 3),
    "C3H8_Qin2000": (70, 463, {0: 420, 1: 6, 2: 1, 3: 36}, 0, 4),
}


def names_of(mech, slots):
    return sorted(mech.names[s] for s in slots)


def states(field, k):
    idx = np.linspace(0, field.n_cells - 1, k).round().astype(int)
    return [(field.p[c], np.r_[field.Y[:, c], field.T[c]]) for c in idx]


@pytest.mark.parametrize("name", TUTORIAL_CASES)
def test_fixture_mechanism_contents(tutorial_case, name):
    mech, field = tutorial_case(name)
    ns, nr, forms, n_noneq, most = EXPECTED[name]
    assert (mech.n_species, mech.n_reactions) == (ns, nr)
    got = {}
    for r in mech.reactions:
        got[r.rtype] = got.get(r.rtype, 0) + 1
    assert got == forms
    assert sum(bool(r.rev_arr) for r in mech.reactions) == n_noneq
    assert max(len(r.rhs) for r in mech.reactions) == most
    assert max(len(r.lhs) for r in mech.reactions) <= M.MAX_STOICH
    assert field.Y.shape == (ns, field.n_cells)
    assert np.abs(field.Y.sum(axis=0) - 1).max() < 1e-5


def test_loader_reads_the_tutorial_rows(tutorial_case):
    """Rows checked against the text of the tutorial files."""
    mech, _ = tutorial_case("NH3_Zhang_counterflow")
    plog = [r for r in mech.reactions if r.rtype == M.RXN_PLOG]
    assert len(plog) == 1                         # un-named-reaction-176, "HNO = H + NO"
    r = plog[0]
    assert names_of(mech, r.lhs) == ["HNO"] and names_of(mech, r.rhs) == ["H", "NO"] and r.reversible
    np.testing.assert_array_equal(r.plog[0], [1e4, 1e5, 1e6, 1e7, 1e8])
    np.testing.assert_array_equal(r.plog[1], [2.0121e16, 1.8259e17, 1.2762e18, 5.6445e18, 9.7111e18])
    np.testing.assert_array_equal(r.plog[2], [-3.021, -3.008, -2.959, -2.855, -2.642])
    np.testing.assert_array_equal(r.plog[3], [24051.2, 24095.4, 24206.2, 24386.9, 24628.9])

    mech, _ = tutorial_case("NC12H26_Yao_counterflow")
    long = [r for r in mech.reactions if len(r.rhs) > 3]
    assert len(long) == 1                         # "OC12H23OOH = 3C2H4 + C2H5 + 2CH2CHO + OH"
    r = long[0]
    assert names_of(mech, r.lhs) == ["OC12H23OOH"] and not r.reversible
    assert names_of(mech, r.rhs) == sorted(["C2H4"] * 3 + ["C2H5"] + ["CH2CHO"] * 2 + ["OH"])

    mech, _ = tutorial_case("H2_Connaire2004")
    h2 = mech.names.index("h2"), mech.names.index("h")
    r = next(r for r in mech.reactions if r.lhs == [h2[0]] and r.rhs == [h2[1], h2[1]])
    assert r.rtype == M.RXN_THIRD_BODY and r.rev_arr      # nonEquilibriumReversibleThirdBodyArrhenius
    assert (r.A, r.beta, r.Ta) == (4.577e16, -1.4, 52533.16533)


@pytest.mark.parametrize("name", TUTORIAL_CASES)
def test_rates_conserve_mass(tutorial_case, name):
    """sum_i dY_i/dt = sum_i W_i omega_i / rho = 0 to round-off at every
    sampled state (the reactions balance exactly with these weights)."""
    mech, field = tutorial_case(name)
    for r in mech.reactions:
        nu = np.zeros(mech.n_species)
        np.add.at(nu, r.lhs, -1.0)
        np.add.at(nu, r.rhs, 1.0)
        assert abs(nu @ mech.W) <= 1e-14 * (np.abs(nu) @ mech.W)
    o = Oracle(mech)
    for p, y in states(field, 16):
        d, _ = o.derivatives(p, y)
        assert abs(d[:-1].sum()) <= 1e-12 * np.abs(d[:-1]).sum()


def long_reactions(mech):
    return [r for r in mech.reactions if len(r.rhs) > 3]


def single(mech, rx):
    m = copy.copy(mech)
    m.reactions = [rx]
    return m


@pytest.mark.parametrize("name", ["Kerosene_Zettervall2020", "NC12H26_Gao2014", "NC12H26_Yao_counterflow"])
def test_long_product_side_closed_form(tutorial_case, name):
    """One irreversible reaction with 4-7 product slots alone: omega_i =
    nu_i A T^beta exp(-Ta/T) prod C_reactants, for every species."""
    mech, _ = tutorial_case(name)
    rng = np.random.default_rng(3)
    T, p = 1500.0, 2e6
    C = rng.uniform(0.1, 1.0, mech.n_species) * 1e-3
    for rx in long_reactions(mech):
        assert not rx.reversible and rx.rtype == M.RXN_ARRHENIUS
        nu = np.zeros(mech.n_species)
        np.add.at(nu, rx.lhs, -1.0)
        np.add.at(nu, rx.rhs, 1.0)
        q = rx.A * T ** rx.beta * np.exp(-rx.Ta / T) * np.prod(C[rx.lhs])
        om = Oracle(single(mech, rx)).rates(p, T, C)
        np.testing.assert_allclose(om, nu * q, rtol=1e-13, atol=1e-13 * abs(q))


def test_long_noneq_reaction_has_no_reverse_rate(tutorial_case):
    """The reference's general path gives a nonEquilibrium reaction with more
    than 3 slots on a side no reverse rate (RFG and JFG take
    isIrreversible == 2 as irreversible); with 3 product slots the same
    reaction keeps its reverse Arrhenius."""
    mech, _ = tutorial_case("Kerosene_Zettervall2020")
    rx = long_reactions(mech)[0]
    T, p = 1500.0, 2e6
    C = np.random.default_rng(4).uniform(0.1, 1.0, mech.n_species) * 1e-3
    irr = Oracle(single(mech, rx)).rates(p, T, C)
    ne = copy.copy(rx)
    ne.reversible, ne.rev_arr = True, (1e12, 0.0, 1000.0)
    np.testing.assert_array_equal(Oracle(single(mech, ne)).rates(p, T, C), irr)
    short, short_ne = copy.copy(rx), copy.copy(ne)
    short.rhs = short_ne.rhs = rx.rhs[:3]
    kr = 1e12 * np.exp(-1000.0 / T) * np.prod(C[rx.rhs[:3]])
    d = Oracle(single(mech, short_ne)).rates(p, T, C) - Oracle(single(mech, short)).rates(p, T, C)
    nu = np.zeros(mech.n_species)
    np.add.at(nu, short.lhs, -1.0)
    np.add.at(nu, short.rhs, 1.0)
    np.testing.assert_allclose(d, -nu * kr, rtol=1e-12, atol=1e-12 * kr)


def g_over_rt(mech, i, T):
    """G/RT of species i from its NASA-7 fit, H and S evaluated separately."""
    low = T < mech.T_common.min() or (T <= mech.T_common.max() and T < mech.T_common[i])
    a = (mech.coeffs_low if low else mech.coeffs_high)[i]
    h = a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5 + a[5] / T
    s = a[0] * np.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4 + a[6]
    return h - s


def test_long_kc_reversible_closed_form(tutorial_case):
    """C3H8_Qin2000's Kc-reversible reaction with 4 product slots
    (C6H5O + O = HCO + 2 C2H2 + CO) alone, on the general path: omega_i =
    nu_i (kf prod C_reactants - kf/Kc prod C_products) over every slot, Kc =
    exp(-dG/RT) (Pstd/(Ru T))^dn from the NASA-7 fits, floored at RFG's
    1.49011611938476e-8 (ReactionFunctionI.H:648-720)."""
    mech, _ = tutorial_case("C3H8_Qin2000")
    (rx,) = long_reactions(mech)
    assert rx.reversible and not rx.rev_arr and len(rx.lhs) == 2 and len(rx.rhs) == 4
    rng = np.random.default_rng(5)
    o = Oracle(single(mech, rx))
    for T, p in ((900.0, 1e5), (1500.0, 2e6), (2400.0, 5e6)):
        C = rng.uniform(0.1, 1.0, mech.n_species) * 1e-3
        dG = sum(g_over_rt(mech, i, T) for i in rx.rhs) - sum(g_over_rt(mech, i, T) for i in rx.lhs)
        Kc = np.exp(-dG) * (M.PSTD / (M.RU * T)) ** (len(rx.rhs) - len(rx.lhs))
        kf = rx.A * T ** rx.beta * np.exp(-rx.Ta / T)
        q = kf * np.prod(C[rx.lhs]) - kf / max(Kc, 1.49011611938476e-8) * np.prod(C[rx.rhs])
        nu = np.zeros(mech.n_species)
        np.add.at(nu, rx.lhs, -1.0)
        np.add.at(nu, rx.rhs, 1.0)
        np.testing.assert_allclose(o.rates(p, T, C), nu * q, rtol=1e-11, atol=1e-11 * kf * np.prod(C[rx.lhs]))


def test_loader_reads_long_kc_reversible_sides():
    text = """
species 5 (A B C D E);
reactions
{
    r0 { type reversibleArrhenius; reaction "A = B + C + D + E"; A 1e10; beta 0; Ta 1000; }
}
"""
    thermo = {n: {"specie": {"molWeight": 10.0},
                  "thermodynamics": {"Tlow": 200, "Thigh": 3500, "Tcommon": 1000,
                                     "highCpCoeffs": [3.0, 0, 0, 0, 0, 0, 0],
                                     "lowCpCoeffs": [3.0, 0, 0, 0, 0, 0, 0]}} for n in "ABCDE"}
    for typ in ("reversibleArrhenius", "irreversibleArrhenius"):
        mech = foam.mechanism(foam.parse(text.replace("reversibleArrhenius", typ)), thermo)
        assert mech.reactions[0].rhs == [1, 2, 3, 4]
        assert mech.reactions[0].reversible == (typ == "reversibleArrhenius")
        a = mech.arrays()
        assert a["slot_stride"] == 4 and list(a["rxn_rhs"]) == [1, 2, 3, 4]
        assert M.Mechanism.from_arrays(a).reactions[0].rhs == [1, 2, 3, 4]
    with pytest.raises(ValueError, match="reactant slots"):
        foam.mechanism(foam.parse(text.replace('"A = B + C + D + E"', '"B + C + D + E = A"')), thermo)


def fd_jacobian(o, p, y, rel=1e-6):
    n = y.size
    J = np.zeros((n, n))
    for j in range(n):
        h = rel * max(abs(y[j]), 1e-8) if j < n - 1 else 1e-3
        yp, ym = y.copy(), y.copy()
        yp[j] += h
        ym[j] -= h
        J[:, j] = (o.derivatives(p, yp)[0] - o.derivatives(p, ym)[0]) / (2 * h)
    return J


@pytest.mark.parametrize("name", ["Kerosene_Zettervall2020", "NC12H26_Gao2014", "H2_Connaire2004", "C3H8_Qin2000"])
def test_exact_jacobian_matches_finite_differences(tutorial_case, name):
    """The exact Jacobian on real mechanisms (long product sides,
    nonEquilibrium rows) against central differences of the RHS, at states
    where every mass fraction is well above zero (the RHS clips negative
    ones, which central differences would straddle)."""
    mech, field = tutorial_case(name)
    o = Oracle(mech)
    for p, y in states(field, 3):
        y = y.copy()
        y[:-1] = np.maximum(y[:-1], 1e-6)
        f, J, _ = o.jacobian(p, y, "exact")
        Jfd = fd_jacobian(o, p, y)
        scale = np.abs(Jfd).max(axis=1, keepdims=True) + 1e-300
        assert (np.abs(J - Jfd) / scale).max() < 1e-4
