"""Pins the oracle's calculate(), tc() and Qdot(): RR against the RHS
(ideal-gas density), mass conservation, closed forms of the time scale for
one reaction, its 1/k scaling, Kc-reversible and nonEquilibrium handling,
and Hf as the NASA-7 enthalpy at 298.15 K."""
import numpy as np
import pytest

import fcs.mechanism as M
from oracle import Oracle

from test_oracle_thermo_rates import R, h_over_rt, tiny_mech

RU = M.RU


@pytest.mark.parametrize("cfg", [0, 1])
def test_calculate_matches_rhs_at_ideal_gas_density(cfg):
    """With rho = p / (R T sum Y/W) and Y >= 0, calculate() is rho dY/dt of
    the RHS, and sum_i RR_i = 0."""
    m = M.config_mechanism(cfg)
    o = Oracle(m)
    fld = M.config_field(cfg, m, n_cells=8)
    for c in range(4):
        Y, T, p = fld.Y[:, c], fld.T[c], fld.p[c]
        rho = p / (RU * T * (Y / m.W).sum())
        rr = o.calculate(rho, T, p, Y)
        d, _ = o.derivatives(p, np.r_[Y, T])
        np.testing.assert_allclose(rr, rho * d[:-1], rtol=1e-10, atol=1e-12 * np.abs(rr).max())
        assert abs(rr.sum()) <= 1e-9 * np.abs(rr).sum()


def test_calculate_does_not_clip():
    """Negative mass fractions enter the concentrations as they are."""
    m = tiny_mech([R(M.RXN_ARRHENIUS, [0], [1], A=1e4, beta=0.0, Ta=1000.0)])
    o = Oracle(m)
    rr = o.calculate(1.0, 1200.0, 1e5, np.array([-0.1, 0.5, 0.3, 0.3]))
    k = 1e4 * np.exp(-1000.0 / 1200.0)
    np.testing.assert_allclose(rr[:2], [-k * (-0.1), k * (-0.1) * m.W[1] / m.W[0]], rtol=1e-13)


def test_tc_single_irreversible_reaction():
    """A + B -> C: sumW = wf, sumWRateByCTot = wf^2 with wf = kf C_A C_B
    (one product slot), so tc = C_tot / (kf C_A C_B)."""
    m = tiny_mech([R(M.RXN_ARRHENIUS, [0, 1], [2], A=3e6, beta=0.4, Ta=6000.0)])
    o = Oracle(m)
    Y = np.array([0.3, 0.2, 0.4, 0.1])
    rho, T = 0.8, 1500.0
    C = rho * Y / m.W
    kf = 3e6 * T ** 0.4 * np.exp(-6000.0 / T)
    assert o.tc(rho, T, 1e5, Y) == pytest.approx(C.sum() / (kf * C[0] * C[1]), rel=1e-13)
    # two product slots double wf: tc halves
    m2 = tiny_mech([R(M.RXN_ARRHENIUS, [0, 1], [2, 3], A=3e6, beta=0.4, Ta=6000.0)])
    assert Oracle(m2).tc(rho, T, 1e5, Y) == pytest.approx(0.5 * C.sum() / (kf * C[0] * C[1]), rel=1e-13)


def test_tc_scales_inversely_with_rates_and_is_vgreat_when_frozen():
    rx = [R(M.RXN_ARRHENIUS, [0], [1], A=2e5, beta=0.0, Ta=3000.0),
          R(M.RXN_ARRHENIUS, [1, 2], [3], A=7e7, beta=-0.5, Ta=9000.0)]
    m = tiny_mech(rx)
    Y = np.array([0.25, 0.25, 0.25, 0.25])
    t1 = Oracle(m).tc(1.1, 1400.0, 1e5, Y)
    for r in m.reactions:
        r.A *= 1e3
    assert Oracle(m).tc(1.1, 1400.0, 1e5, Y) == pytest.approx(t1 * 1e-3, rel=1e-12)
    assert Oracle(m).tc(1.1, 1400.0, 1e5, np.array([0.0, 0.0, 0.5, 0.5])) == 1e300


def test_tc_reversible_counts_reverse_rate_and_noneq_does_not():
    """A reversible reaction adds wr = n_lhs kr C_R (kr = kf / Kc); a
    nonEquilibrium one adds nothing (Tc.C:583 tests isIrreversible == 0)."""
    Y = np.array([0.3, 0.2, 0.4, 0.1])
    rho, T, p = 0.8, 1500.0, 1e5
    base = dict(A=3e6, beta=0.4, Ta=6000.0)
    irr = tiny_mech([R(M.RXN_ARRHENIUS, [0, 1], [2], **base)])
    rev = tiny_mech([R(M.RXN_ARRHENIUS, [0, 1], [2], rev=True, **base)])
    neq = tiny_mech([R(M.RXN_ARRHENIUS, [0, 1], [2], rev=True, **base)])
    neq.reactions[0].rev_arr = (1e9, 0.0, 100.0)
    C = rho * Y / irr.W
    o = Oracle(rev)
    # kr from the oracle's own rates: q = kf CF - kr CR
    q = -o.rates(p, T, C)[0]
    kf = 3e6 * T ** 0.4 * np.exp(-6000.0 / T)
    kr = (kf * C[0] * C[1] - q) / C[2]
    wf, wr = kf * C[0] * C[1], 2 * kr * C[2]
    assert o.tc(rho, T, p, Y) == pytest.approx((wf + wr) / (wf * wf + wr * wr) * C.sum(), rel=1e-9)
    assert Oracle(neq).tc(rho, T, p, Y) == Oracle(irr).tc(rho, T, p, Y)


def test_hf_is_nasa7_enthalpy_at_298_15(mech1):
    o = Oracle(mech1)
    hf = o.hf()
    Ts = 298.15
    for i in range(mech1.n_species):
        expect = RU / mech1.W[i] * Ts * h_over_rt(mech1.coeffs_low[i], Ts)
        assert hf[i] == pytest.approx(expect, rel=1e-12, abs=1e-6)


def test_qdot_is_minus_hf_weighted_rr(mech1):
    o = Oracle(mech1)
    rng = np.random.default_rng(3)
    rr = rng.standard_normal(mech1.n_species)
    assert o.qdot(rr) == pytest.approx(-(o.hf() * rr).sum(), rel=1e-12)
