"""Pins the oracle's thermo, rate constants and source terms by closed forms,
limits and invariants (no reference test covers them; the reference's
OptReaction needs OpenFOAM to build)."""
import numpy as np
import pytest

import fcs.mechanism as M
from oracle import Oracle

RU = M.RU


def tiny_mech(reactions, ns=4, seed=7):
    """ns species with random NASA-7 fits and the given reactions."""
    rng = np.random.default_rng(seed)
    lows, highs, tcs = [], [], []
    W = np.array([10.0, 20.0, 30.0, 40.0, 50.0, 60.0])[:ns]
    for w in W:
        lo, hi, tc = M._nasa7_for(rng, w)
        lows.append(lo)
        highs.append(hi)
        tcs.append(tc)
    m = M.Mechanism(names=[f"S{i}" for i in range(ns)], W=W, T_low=np.full(ns, 200.0),
                    T_high=np.full(ns, 3500.0), T_common=np.array(tcs), coeffs_low=np.array(lows),
                    coeffs_high=np.array(highs))
    m.reactions = reactions
    return m


def R(rtype, lhs, rhs, A=1e3, beta=0.5, Ta=3000.0, rev=False, **kw):
    return M.Reaction(rtype=rtype, reversible=rev, lhs=lhs, rhs=rhs, A=A, beta=beta, Ta=Ta, **kw)


# ------------------------------------------------------------------ thermo
def h_over_rt(a, T):
    return a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5 + a[5] / T


def s_over_r(a, T):
    return a[0] * np.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4 + a[6]


@pytest.mark.parametrize("T", [300.0, 950.0, 1000.0, 1500.0, 2800.0])
def test_thermo_closed_form(mech1, T):
    o = Oracle(mech1)
    th = o.thermo(T)
    for i in range(mech1.n_species):
        low = T < mech1.T_common.min() or (T <= mech1.T_common.max() and T < mech1.T_common[i])
        a = (mech1.coeffs_low if low else mech1.coeffs_high)[i]
        r = RU / mech1.W[i]
        cp = r * (a[0] + a[1] * T + a[2] * T ** 2 + a[3] * T ** 3 + a[4] * T ** 4)
        assert th["cp"][i] == pytest.approx(cp, rel=1e-13)
        assert th["ha"][i] == pytest.approx(r * T * h_over_rt(a, T), rel=1e-12, abs=1e-6)
        assert th["negG"][i] == pytest.approx(s_over_r(a, T) - h_over_rt(a, T), rel=1e-12, abs=1e-10)


@pytest.mark.parametrize("T", [700.0, 1400.0, 2500.0])
def test_thermo_derivative_identities(mech1, T):
    """dH/dT = cp, d(-G/RT)/dT = H/(R T^2) = dBdT, d cp/dT = dcpdT (central FD)."""
    o = Oracle(mech1)
    h = 1e-3
    tp, tm, t0 = o.thermo(T + h), o.thermo(T - h), o.thermo(T)
    sel = np.abs(mech1.T_common - T) > 1.0   # away from the fit switch
    np.testing.assert_allclose(((tp["ha"] - tm["ha"]) / (2 * h))[sel], t0["cp"][sel], rtol=1e-7)
    np.testing.assert_allclose(((tp["negG"] - tm["negG"]) / (2 * h))[sel], t0["dBdT"][sel], rtol=1e-6)
    np.testing.assert_allclose(((tp["cp"] - tm["cp"]) / (2 * h))[sel], t0["dcpdT"][sel], rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(t0["dBdT"], t0["ha"] * mech1.W / (RU * T * T), rtol=1e-12)


def test_fits_continuous_at_T_common(mech1):
    """The generator's low/high fits agree at T_common (value of cp, H, S)."""
    for i in range(mech1.n_species):
        Tc = mech1.T_common[i]
        lo, hi = mech1.coeffs_low[i], mech1.coeffs_high[i]
        assert h_over_rt(lo, Tc) == pytest.approx(h_over_rt(hi, Tc), rel=1e-12)
        assert s_over_r(lo, Tc) == pytest.approx(s_over_r(hi, Tc), rel=1e-12)


# ------------------------------------------------------------------ rates
def conc(m, p, T, x):
    return np.asarray(x) * p / (RU * T)


def test_arrhenius_closed_form():
    A, b, Ta = 2.5e7, -0.7, 8000.0
    m = tiny_mech([R(M.RXN_ARRHENIUS, [0], [1], A=A, beta=b, Ta=Ta)])
    o = Oracle(m)
    T, p = 1300.0, 1e5
    C = conc(m, p, T, [0.4, 0.3, 0.2, 0.1])
    om = o.rates(p, T, C)
    k = A * T ** b * np.exp(-Ta / T)
    np.testing.assert_allclose(om, [-k * C[0], k * C[0], 0, 0], rtol=1e-13)


def test_second_order_same_species():
    """2 S0 -> S1: d[S0]/dt = -2 k [S0]^2 (stoichiometry expanded into slots)."""
    m = tiny_mech([R(M.RXN_ARRHENIUS, [0, 0], [1], A=5e6, beta=0.0, Ta=1000.0)])
    o = Oracle(m)
    T, p = 1100.0, 2e5
    C = conc(m, p, T, [0.5, 0.5, 0, 0])
    k = 5e6 * np.exp(-1000.0 / T)
    om = o.rates(p, T, C)
    np.testing.assert_allclose(om[:2], [-2 * k * C[0] ** 2, k * C[0] ** 2], rtol=1e-13)


def test_third_body_closed_form():
    eff = np.array([1.0, 2.5, 0.5, 3.0])
    m = tiny_mech([R(M.RXN_THIRD_BODY, [0, 1], [2], A=1e5, beta=0.2, Ta=500.0, eff=eff)])
    o = Oracle(m)
    T, p = 1500.0, 1e5
    C = conc(m, p, T, [0.25, 0.25, 0.25, 0.25])
    k = 1e5 * T ** 0.2 * np.exp(-500.0 / T) * (eff @ C)
    np.testing.assert_allclose(o.rates(p, T, C)[2], k * C[0] * C[1], rtol=1e-13)


def falloff_mech(rtype, troe=(0.0, 1.0, 1.0, 1e30)):
    eff = np.ones(4)
    return tiny_mech([R(rtype, [0], [1, 2], A=2e9, beta=0.0, Ta=20000.0, eff=eff, A0=4e12, beta0=0.0,
                        Ta0=15000.0, troe=troe)])


def test_lindemann_pressure_limits():
    """Pr -> infinity gives k_inf, Pr -> 0 gives k0*M."""
    m = falloff_mech(M.RXN_LINDEMANN)
    o = Oracle(m)
    T = 1400.0
    kinf = 2e9 * np.exp(-20000.0 / T)
    k0 = 4e12 * np.exp(-15000.0 / T)
    for p, lim in ((1e40, "high"), (1e-40, "low")):
        C = conc(m, p, T, [1.0, 0, 0, 0])
        kf = -o.rates(p, T, C)[0] / C[0]
        target = kinf if lim == "high" else k0 * C.sum()
        assert kf == pytest.approx(target, rel=1e-12)


def test_troe_broadening_bounds():
    """log F = log Fcent / (1 + x^2) lies between log Fcent and 0 at every
    pressure, so k_Lindemann * Fcent <= k_Troe <= k_Lindemann."""
    troe = (0.6, 200.0, 1500.0, 5000.0)
    ot, ol = Oracle(falloff_mech(M.RXN_TROE, troe)), Oracle(falloff_mech(M.RXN_LINDEMANN))
    T = 1400.0
    alpha, Tsss, Ts, Tss = troe
    Fc = (1 - alpha) * np.exp(-T / Tsss) + alpha * np.exp(-T / Ts) + np.exp(-Tss / T)
    m = falloff_mech(M.RXN_TROE, troe)
    for p in np.logspace(-6, 14, 21):
        C = conc(m, p, T, [1.0, 0, 0, 0])
        kt = -ot.rates(p, T, C)[0]
        kl = -ol.rates(p, T, C)[0]
        assert kl * Fc * (1 - 1e-12) <= kt <= kl * (1 + 1e-12)


def test_troe_reduces_to_lindemann_when_Fcent_is_one():
    """alpha = 0, T*** huge and T** huge make Fcent = 1, so F = 1."""
    lind = Oracle(falloff_mech(M.RXN_LINDEMANN))
    troe = Oracle(falloff_mech(M.RXN_TROE, troe=(0.0, 1e300, 1.0, 1e300)))
    T, p = 1300.0, 3e4
    m = falloff_mech(M.RXN_LINDEMANN)
    C = conc(m, p, T, [0.7, 0.1, 0.1, 0.1])
    np.testing.assert_allclose(troe.rates(p, T, C), lind.rates(p, T, C), rtol=1e-13)


def test_troe_center_point():
    """At log10(Pr) = -c the broadening factor is exactly Fcent
    (Gilbert, Luther & Troe 1983)."""
    alpha, Tsss, Ts, Tss = 0.6, 200.0, 1500.0, 5000.0
    m = falloff_mech(M.RXN_TROE, troe=(alpha, Tsss, Ts, Tss))
    o = Oracle(m)
    T = 1400.0
    Fc = (1 - alpha) * np.exp(-T / Tsss) + alpha * np.exp(-T / Ts) + np.exp(-Tss / T)
    c = -0.4 - 0.67 * np.log10(Fc)
    kinf = 2e9 * np.exp(-20000.0 / T)
    k0 = 4e12 * np.exp(-15000.0 / T)
    Mc = 10 ** (-c) * kinf / k0          # Pr = k0 M / kinf = 10^-c
    p = Mc * RU * T
    C = conc(m, p, T, [1.0, 0, 0, 0])
    Pr = k0 * Mc / kinf
    kf = -o.rates(p, T, C)[0] / C[0]
    assert kf == pytest.approx(kinf * Pr / (1 + Pr) * Fc, rel=1e-10)


def plog_mech():
    P = np.array([1e4, 1e5, 1e6, 1e7, 1e8])
    A = np.array([1e6, 3e6, 2e7, 5e7, 1e8])
    b = np.array([0.1, 0.0, -0.2, 0.3, 0.5])
    Ta = np.array([5000.0, 6000.0, 7000.0, 6500.0, 8000.0])
    return tiny_mech([R(M.RXN_PLOG, [0], [1], A=1.0, beta=0, Ta=0, plog=(P, A, b, Ta))]), (P, A, b, Ta)


def test_plog_at_tabulated_pressures_and_midpoints():
    m, (P, A, b, Ta) = plog_mech()
    o = Oracle(m)
    T = 1250.0
    lnk = np.log(A) + b * np.log(T) - Ta / T
    for j in range(5):
        C = conc(m, P[j], T, [1.0, 0, 0, 0])
        assert -o.rates(P[j], T, C)[0] / C[0] == pytest.approx(np.exp(lnk[j]), rel=1e-12)
    for j in range(4):
        pm = np.sqrt(P[j] * P[j + 1])   # geometric mean: log k is the arithmetic mean
        C = conc(m, pm, T, [1.0, 0, 0, 0])
        assert -o.rates(pm, T, C)[0] / C[0] == pytest.approx(np.exp(0.5 * (lnk[j] + lnk[j + 1])), rel=1e-12)
    for p, j in ((1e2, 0), (1e10, 4)):  # clamped to the end rows
        C = conc(m, p, T, [1.0, 0, 0, 0])
        assert -o.rates(p, T, C)[0] / C[0] == pytest.approx(np.exp(lnk[j]), rel=1e-12)


def kc_independent(m, T, lhs, rhs):
    """Kc = exp(-dG/RT) (Pstd/(Ru T))^dn from H and S evaluated separately."""
    def g_over_rt(i):
        low = T < m.T_common.min() or (T <= m.T_common.max() and T < m.T_common[i])
        a = (m.coeffs_low if low else m.coeffs_high)[i]
        return h_over_rt(a, T) - s_over_r(a, T)
    dG = sum(g_over_rt(i) for i in rhs) - sum(g_over_rt(i) for i in lhs)
    return np.exp(-dG) * (M.PSTD / (RU * T)) ** (len(rhs) - len(lhs))


@pytest.mark.parametrize("lhs,rhs", [([0], [1]), ([0], [1, 2]), ([0, 1], [2]), ([0, 1], [2, 3]),
                                     # product sides of 4 and 5 slots: the general path (RFG)
                                     ([0], [1, 2, 3, 1]), ([0, 1], [2, 3, 2, 2, 2])])
def test_detailed_balance(lhs, rhs):
    """q = 0 at a composition satisfying prod C_rhs / prod C_lhs = Kc."""
    T, p = 1800.0, 1e5
    # a thermo draw with Kc well above the Kc floor (the long sides' random
    # thermo draws give Kc of 1e8 and more: no upper bound there)
    hi = 1e4 if len(rhs) <= 3 else np.inf
    for seed in range(11, 200):
        m = tiny_mech([R(M.RXN_ARRHENIUS, lhs, rhs, A=1e8, beta=0.3, Ta=4000.0, rev=True)], seed=seed)
        Kc = kc_independent(m, T, lhs, rhs)
        if 1e-4 < Kc < hi:
            break
    assert 1e-4 < Kc < hi
    o = Oracle(m)
    C = np.full(4, 0.01)
    # scale the first product of a single slot (and no reactant slot) so that
    # the mass-action ratio equals Kc
    ratio = np.prod(C[rhs]) / np.prod(C[lhs])
    C[next(s for s in rhs if rhs.count(s) == 1 and s not in lhs)] *= Kc / ratio
    om = o.rates(p, T, C)
    kf = 1e8 * T ** 0.3 * np.exp(-4000.0 / T)
    scale = kf * np.prod(C[lhs])
    assert abs(om[lhs[0]]) < 1e-10 * scale
    # away from equilibrium the reverse rate is kf/Kc
    C2 = np.full(4, 0.01)
    q = -o.rates(p, T, C2)[lhs[0]] / lhs.count(lhs[0])
    q_expected = kf * np.prod(C2[lhs]) - kf / max(Kc, 1.4901171103413047e-8) * np.prod(C2[rhs])
    assert q == pytest.approx(q_expected, rel=1e-10)


def test_kc_floor():
    """Kc below 1.4901171103413047e-8 is clamped (ReactionFunctionI.H:17)."""
    m = tiny_mech([R(M.RXN_ARRHENIUS, [0], [1], A=1e5, beta=0.0, Ta=1000.0, rev=True)], seed=3)
    # make S1 very unstable so that Kc << floor
    m.coeffs_low[1, 5] += 1e5
    m.coeffs_high[1, 5] += 1e5
    o = Oracle(m)
    T, p = 1000.0, 1e5
    assert kc_independent(m, T, [0], [1]) < 1e-20
    C = np.array([0.0, 0.02, 0.0, 0.0])
    kf = 1e5 * np.exp(-1000.0 / T)
    assert o.rates(p, T, C)[0] == pytest.approx(kf / 1.4901171103413047e-8 * C[1], rel=1e-12)


def test_kc_floor_of_the_general_path():
    """A Kc-reversible reaction with more than 3 product slots takes the
    general path, whose RHS floors Kc at 1.49011611938476e-8
    (ReactionFunctionI.H:688) while its Jacobian floors it at
    1.4901171103413047e-8 (JacobianFunctionI.H:3005); with no reactant
    present, the product's rate is the reverse rate alone in both, so the
    RHS and the Jacobian evaluation's f(y) differ by the ratio of the floors."""
    m = tiny_mech([R(M.RXN_ARRHENIUS, [0], [1, 2, 3, 1], A=1e5, beta=0.0, Ta=1000.0, rev=True)], seed=3)
    m.coeffs_low[1, 5] += 1e5    # S1 very unstable: Kc << floor
    m.coeffs_high[1, 5] += 1e5
    o = Oracle(m)
    T, p = 1000.0, 1e5
    assert kc_independent(m, T, [0], [1, 2, 3, 1]) < 1e-20
    C = np.array([0.0, 0.02, 0.03, 0.01])
    kf = 1e5 * np.exp(-1000.0 / T)
    q = -kf / 1.49011611938476e-08 * C[1] * C[2] * C[3] * C[1]
    np.testing.assert_allclose(o.rates(p, T, C), q * np.array([-1.0, 2.0, 1.0, 1.0]), rtol=1e-12)
    # f(y) of the RHS and of the Jacobian evaluation on the same state
    Y = C * m.W / (C * m.W).sum()
    y = np.r_[Y, T]
    f_rhs, _ = o.derivatives(p, y)
    f_jac, _, _ = o.jacobian(p, y)
    np.testing.assert_allclose(f_jac[1:4] / f_rhs[1:4], 1.49011611938476e-08 / 1.4901171103413047e-8, rtol=1e-12)


# ------------------------------------------------------------------ invariants
@pytest.mark.parametrize("cfg", [0, 1])
def test_element_and_mass_conservation(cfg):
    m = M.config_mechanism(cfg)
    o = Oracle(m)
    fld = M.config_field(cfg, m, n_cells=8)
    for c in range(8):
        C = conc(m, fld.p[c], fld.T[c], fld.Y[:, c] / m.W / (fld.Y[:, c] / m.W).sum())
        om = o.rates(fld.p[c], fld.T[c], C)
        scale = np.abs(om).max() * 10
        np.testing.assert_allclose(m.composition.T @ om, 0, atol=1e-12 * scale)
        assert abs(m.W @ om) < 1e-12 * scale * m.W.max()


@pytest.mark.parametrize("cfg", [0, 1])
def test_derivatives_conserve_mass_and_enthalpy(cfg):
    m = M.config_mechanism(cfg)
    o = Oracle(m)
    fld = M.config_field(cfg, m, n_cells=8)
    for c in range(8):
        y = np.concatenate([fld.Y[:, c], [fld.T[c]]])
        d, yc = o.derivatives(fld.p[c], y)
        th = o.thermo(fld.T[c])
        sY = np.abs(d[:-1]).max()
        assert abs(d[:-1].sum()) < 1e-12 * sY
        # d/dt sum(Y h) = sum(dY h) + sum(Y cp) dT/dt = 0
        lhs = d[:-1] @ th["ha"] + (yc[:-1] @ th["cp"]) * d[-1]
        assert abs(lhs) < 1e-11 * np.abs(d[:-1] * th["ha"]).sum()


def test_derivatives_clip_negative_mass_fractions(mech0):
    o = Oracle(mech0)
    y = np.concatenate([np.full(mech0.n_species, 0.1), [1200.0]])
    y[2] = -0.05
    _, yc = o.derivatives(1e5, y)
    assert yc[2] == 0.0 and yc[-1] == 1200.0
