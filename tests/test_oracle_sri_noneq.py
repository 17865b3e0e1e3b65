"""Pins the oracle's SRI falloff and nonEquilibrium reverse rates by special
points of the SRI blending function, the falloff limits, independence of the
nonEquilibrium reverse rate from the thermodynamic data, central finite
differences of the Jacobian, and the dictionary loader's reading of both
forms."""
import numpy as np
import pytest

import fcs.mechanism as M
from fcs import foam
from oracle import Oracle

from test_oracle_jacobian import fd_jacobian
from test_oracle_thermo_rates import R, tiny_mech

RU = M.RU
SRI = (0.6, 400.0, 1500.0, 1.1, 0.2)     # a, b, c, d, e


def sri_mech(sri=SRI, beta0=0.0, Ta0=15000.0):
    return tiny_mech([R(M.RXN_SRI, [0], [1, 2], A=2e9, beta=0.0, Ta=20000.0, eff=np.ones(4),
                        A0=4e12, beta0=beta0, Ta0=Ta0, sri=sri)])


def kf_of(o, p, T, C):
    return -o.rates(p, T, C)[0] / C[0]


@pytest.mark.parametrize("Pr,X", [(1.0, 1.0), (10.0, 0.5), (0.01, 0.2), (1e3, 0.1)])
def test_sri_blending_at_exact_points(Pr, X):
    """log10 Pr = 0, 1, -2, 3 give X = 1/(1 + log10(Pr)^2) = 1, 1/2, 1/5,
    1/10, so kf = k_inf Pr/(1+Pr) d psi^X T^e there (Stewart et al. 1989)."""
    o = Oracle(sri_mech())
    T, p = 1400.0, 1e5
    kinf = 2e9 * np.exp(-20000.0 / T)
    k0 = 4e12 * np.exp(-15000.0 / T)
    Mc = Pr * kinf / k0
    C = np.array([0.6, 0.2, 0.1, 0.1]) * Mc    # unit efficiencies: M = sum C
    a, b, c, d, e = SRI
    psi = a * np.exp(-b / T) + np.exp(-T / c)
    expect = kinf * Pr / (1 + Pr) * d * psi ** X * T ** e
    assert kf_of(o, p, T, C) == pytest.approx(expect, rel=1e-12)


def test_sri_pressure_limits():
    """High pressure: k_inf d T^e (X -> 0); low pressure: k0 M d T^e psi^X
    with X at the log10 clamp of Pr at OpenFOAM's small (1e-15)."""
    o = Oracle(sri_mech())
    T = 1400.0
    kinf = 2e9 * np.exp(-20000.0 / T)
    k0 = 4e12 * np.exp(-15000.0 / T)
    a, b, c, d, e = SRI
    psi = a * np.exp(-b / T) + np.exp(-T / c)
    C = np.array([1.0, 0, 0, 0]) * 1e40
    assert kf_of(o, 1e5, T, C) == pytest.approx(kinf * d * T ** e, rel=2e-3)
    C = np.array([1.0, 0, 0, 0]) * 1e-30
    Xc = 1.0 / (1.0 + 15.0 ** 2)
    assert kf_of(o, 1e5, T, C) == pytest.approx(k0 * C.sum() * d * psi ** Xc * T ** e, rel=1e-9)


def test_sri_reduces_to_lindemann_scaled():
    """a = 0, c -> infinity, e = 0: psi = 1, so F = d at every pressure."""
    d = 1.7
    o = Oracle(sri_mech(sri=(0.0, 1.0, 1e300, d, 0.0)))
    lind = Oracle(tiny_mech([R(M.RXN_LINDEMANN, [0], [1, 2], A=2e9, beta=0.0, Ta=20000.0, eff=np.ones(4),
                               A0=4e12, beta0=0.0, Ta0=15000.0)]))
    for p in (1e2, 1e5, 1e9):
        C = np.array([0.7, 0.1, 0.1, 0.1]) * p / (RU * 1300.0)
        assert kf_of(o, p, 1300.0, C) == pytest.approx(d * kf_of(lind, p, 1300.0, C), rel=1e-13)


def noneq_mech(tb=False, seed=7):
    rt = M.RXN_THIRD_BODY if tb else M.RXN_ARRHENIUS
    rx = R(rt, [0, 1], [2, 3], A=3e8, beta=0.3, Ta=9000.0, rev=True, eff=np.array([1.0, 2.0, 0.5, 1.5]))
    rx.rev_arr = (7e7, -0.4, 4000.0)
    return tiny_mech([rx], seed=seed)


@pytest.mark.parametrize("tb", [False, True])
def test_noneq_reverse_rate_is_its_own_arrhenius(tb):
    """q = kf CF - kr CR with kr = A_r T^b_r exp(-Ta_r/T) (times M for the
    third-body form), whatever the thermodynamic data say."""
    T, p = 1600.0, 2e5
    C = np.array([0.3, 0.2, 0.4, 0.1]) * p / (RU * T)
    kf = 3e8 * T ** 0.3 * np.exp(-9000.0 / T)
    kr = 7e7 * T ** -0.4 * np.exp(-4000.0 / T)
    Mc = np.array([1.0, 2.0, 0.5, 1.5]) @ C if tb else 1.0
    q = Mc * (kf * C[0] * C[1] - kr * C[2] * C[3])
    for seed in (7, 8):   # different NASA-7 fits: no Kc enters
        om = Oracle(noneq_mech(tb, seed)).rates(p, T, C)
        np.testing.assert_allclose(om, [-q, -q, q, q], rtol=1e-12)


def mixed_mech():
    """SRI with dPr/dT = 0 (k0 and k_inf share beta and Ta, so the
    reference's dF/dT without the dF/dPr dPr/dT term is the full
    derivative) plus nonEquilibrium Arrhenius and third-body reactions."""
    r1 = R(M.RXN_SRI, [0], [1, 2], A=2e9, beta=0.1, Ta=12000.0, eff=np.array([1.0, 2.0, 0.7, 1.3]),
           A0=4e10, beta0=0.1, Ta0=12000.0, sri=SRI, rev=True)
    r2 = R(M.RXN_ARRHENIUS, [0, 1], [2, 3], A=3e8, beta=0.3, Ta=9000.0, rev=True)
    r2.rev_arr = (7e7, -0.4, 4000.0)
    r3 = R(M.RXN_THIRD_BODY, [1, 2], [0], A=5e6, beta=0.0, Ta=3000.0, rev=True,
           eff=np.array([1.0, 2.0, 0.5, 1.5]))
    r3.rev_arr = (2e9, 0.2, 11000.0)
    m = tiny_mech([r1, r2, r3])
    m.W = np.array([60.0, 20.0, 40.0, 40.0])   # every reaction mass-balanced
    return m


def test_sri_noneq_jacobian_matches_finite_differences():
    m = mixed_mech()
    o = Oracle(m)
    for T in (1100.0, 1700.0):
        y = np.array([0.4, 0.2, 0.3, 0.1, T])
        f, J, _ = o.jacobian(1e5, y, "exact")
        Jfd = fd_jacobian(o, 1e5, y)
        scale = np.abs(Jfd).max(axis=1, keepdims=True) + 1e-300
        assert (np.abs(J - Jfd) / scale).max() < 2e-5
        np.testing.assert_allclose(f, o.derivatives(1e5, y)[0], rtol=1e-12, atol=1e-12 * np.abs(f).max())


def test_sri_jacobian_omits_the_pr_chain_term():
    """With dPr/dT != 0 the T column differs from central differences by
    exactly the dropped k0/(1+Pr) M dF/dPr dPr/dT CF term, as in the
    reference (Jacobian.C:699-708)."""
    m = sri_mech(beta0=0.0, Ta0=15000.0)
    o = Oracle(m)
    T, p = 1400.0, 1e5
    y = np.array([0.7, 0.1, 0.1, 0.1, T])
    _, _, _, Jc = o.jacobian(p, y, "exact", want_jc=True)
    rho = p / (RU * T * (y[:-1] / m.W).sum())
    C = rho * y[:-1] / m.W
    h = 1e-3
    colT = (o.rates(p, T + h, C) - o.rates(p, T - h, C)) / (2 * h)
    a, b, c, d, e = SRI
    kinf, k0 = 2e9 * np.exp(-20000.0 / T), 4e12 * np.exp(-15000.0 / T)
    Mc = C.sum()
    Pr = k0 * Mc / kinf
    lg = np.log10(Pr)
    X = 1 / (1 + lg * lg)
    psi = a * np.exp(-b / T) + np.exp(-T / c)
    F = d * psi ** X * T ** e
    dFdPr = F * np.log(psi) * (-(X * X) * 2 * lg / (np.log(10) * Pr))
    dPrdT = Pr * (15000.0 - 20000.0) / T ** 2
    missing = k0 / (1 + Pr) * Mc * dFdPr * dPrdT * C[0]
    np.testing.assert_allclose(Jc[:4, 4] - colT, [missing, -missing, -missing, 0.0],
                               rtol=1e-4, atol=1e-6 * np.abs(colT).max())


FOAM = """
species 4 ( A B C D );
reactions
{
    un-named-reaction-0
    {
        type            reversibleArrheniusSRIFallOff;
        reaction        "A = B + C";
        k0 { A 4e10; beta 0.1; Ta 12000; }
        kInf { A 2e9; beta 0.1; Ta 12000; }
        F { a 0.6; b 400; c 1500; d 1.1; e 0.2; }
        thirdBodyEfficiencies { coeffs 4 ( (A 1) (B 2) (C 0.7) (D 1.3) ); }
    }
    un-named-reaction-1
    {
        type            nonEquilibriumReversibleArrhenius;
        reaction        "A + B = C + D";
        forward { A 3e8; beta 0.3; Ta 9000; }
        reverse { A 7e7; beta -0.4; Ta 4000; }
    }
    un-named-reaction-2
    {
        type            nonEquilibriumReversibleThirdBodyArrhenius;
        reaction        "B + C = A";
        forward { A 5e6; beta 0; Ta 3000; coeffs 4 ( (A 1) (B 2) (C 0.5) (D 1.5) ); }
        reverse { A 2e9; beta 0.2; Ta 11000; coeffs 4 ( (A 1) (B 1) (C 1) (D 1) ); }
    }
}
"""


def test_loader_reads_sri_and_noneq():
    ref = mixed_mech()
    thermo = {"species": ["A", "B", "C", "D"]}
    for i, n in enumerate("ABCD"):
        thermo[n] = {"specie": {"molWeight": ref.W[i]},
                     "thermodynamics": {"Tlow": 200, "Thigh": 3500, "Tcommon": ref.T_common[i],
                                        "highCpCoeffs": list(ref.coeffs_high[i]),
                                        "lowCpCoeffs": list(ref.coeffs_low[i])}}
    m = foam.mechanism(foam.parse(FOAM), thermo)
    a, b = m.arrays(), ref.arrays()
    for k in ("rxn_type", "rxn_reversible", "rxn_lhs", "rxn_rhs", "A", "beta", "Ta", "A0", "beta0", "Ta0",
              "sri", "A_rev", "beta_rev", "Ta_rev"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    ne = slice(2 * 4, 3 * 4)   # third-body row: forward efficiencies for both directions
    np.testing.assert_array_equal(a["efficiencies"][ne], [1.0, 2.0, 0.5, 1.5])
    C = np.array([0.3, 0.2, 0.4, 0.1]) * 10.0
    np.testing.assert_array_equal(Oracle(m).rates(1e5, 1500.0, C), Oracle(ref).rates(1e5, 1500.0, C))
