"""Pins the reference's f(y0) stage terms of the Rosenbrock steppers by a
closed form.

The reference's Jacobian evaluation returns f(y0) where a Rosenbrock method
expects df/dt, so stage 1 solves (I/(gamma h) - J) k1 = (1 + d1 h) f(y0) and
every later stage adds d_i h f(y0) (src/ODE/OptRodas34/OptRodas34.C:345-371,
407-417, 477-492, 554-574; OptRosenbrock34.C and OptRodas23.C,
OptRosenbrock23.C likewise).  The order tests switch that term off, and a
Radau comparison cannot see an O(h^2) term per step.  Here one attempt of
every Rosenbrock solver runs on a mechanism whose ODE is exactly linear,
y' = J y with J = [[-k, 0], [k, 0]] on (Y_A, Y_B) and T constant (A -> B,
k temperature-independent, A and B with identical molar mass and thermo),
and is compared with the one-step map of that linear system evaluated in
50-digit arithmetic from the published tableaus in the transformed
(Hairer-Wanner) form the reference codes:

    (I/(gamma h) - J) k_s = J (y0 + sum_j a_sj k_j) + d_s h J y0 + sum_j (c_sj / h) k_j
    y1 = y0 + sum_j b_j k_j,   err = sum_j e_j k_j.

A sign, coefficient or stage-placement mistake in the d_i h f(y0) terms
moves the result by ~d h relative, far above the bar (1e-11 relative); the
same attempt with the terms switched off must match the map with d = 0.
"""
import mpmath as mp
import numpy as np
import pytest

import fcs.mechanism as M
from oracle import Oracle, lib as orc_lib

from test_oracle_thermo_rates import R, tiny_mech

mp.mp.dps = 50
F = mp.mpf

# Rodas3(4), Sandu et al. 1997 (src/ODE/OptRodas34/OptRodas34.H:116-149):
# six stages, stage 6 at y0 + sum a5j k_j + k5, new state y0 + sum a5j k_j +
# k5 + k6, error k6.
_a5 = [F("1.221224509226641"), F("6.019134481288629"), F("12.53708332932087"), F("-0.6878860361058950")]
RODAS34 = dict(
    gamma=F("0.25"),
    A=[[], [F("1.544")], [F("0.9466785280815826"), F("0.2557011698983284")],
       [F("3.314825187068521"), F("2.896124015972201"), F("0.9986419139977817")], _a5, _a5 + [F(1)]],
    C=[[], [F("-5.6688")], [F("-2.430093356833875"), F("-0.2063599157091915")],
       [F("-0.1073529058151375"), F("-9.594562251023355"), F("-20.47028614809616")],
       [F("7.496443313967647"), F("-10.24680431464352"), F("-33.99990352819905"), F("11.70890893206160")],
       [F("8.083246795921522"), F("-7.981132988064893"), F("-31.52159432874371"), F("16.31930543123136"),
        F("-6.058818238834054")]],
    D=[F("0.25"), F("-0.1043"), F("0.1035"), F("-0.3620000000000023e-01"), F(0), F(0)],
    B=_a5 + [F(1), F(1)],
    E=[F(0)] * 5 + [F(1)])

# ROS3(4), Shampine 1982 (src/ODE/OptRosenbrock34/OptRosenbrock34.H:127-158):
# stage 4 reuses stage 3's derivative (its f argument is stage 3's).
_r3 = [F(48) / 25, F(6) / 25]
ROS34 = dict(
    gamma=F(1) / 2,
    A=[[], [F(2)], _r3, _r3],
    C=[[], [F(-8)], [F(372) / 25, F(12) / 5], [F(-112) / 125, F(-54) / 125, F(-2) / 5]],
    D=[F(1) / 2, F(-3) / 2, F(605) / 250, F(29) / 250],
    B=[F(19) / 9, F(1) / 2, F(25) / 108, F(125) / 108],
    E=[F(34) / 108, F(7) / 36, F(0), F(125) / 108])

# Rodas3(2), Sandu et al. 1997 (src/ODE/OptRodas23/OptRodas23.H): stage 2
# takes f(y0), stage 3 f(y0 + 2 k1), stage 4 f(y0 + 2 k1 + k3); new state
# y0 + 2 k1 + k3 + k4, error k4.
RODAS23 = dict(
    gamma=F(1) / 2,
    A=[[], [F(0)], [F(2)], [F(2), F(0), F(1)]],
    C=[[], [F(4)], [F(1), F(-1)], [F(1), F(-1), F(-8) / 3]],
    D=[F(1) / 2, F(3) / 2, F(0), F(0)],
    B=[F(2), F(0), F(1), F(1)],
    E=[F(0), F(0), F(0), F(1)])

# ROS3 with embedded order 2, Sandu et al. 1997 (src/ODE/OptRosenbrock23/
# OptRosenbrock23.H:118-140): stage 3 reuses stage 2's derivative.
_g23 = F("0.43586652150845899941601945119356")
ROS23 = dict(
    gamma=_g23,
    A=[[], [F(1)], [F(1)]],
    C=[[], [F("-1.0156171083877702091975600115545")],
       [F("4.0759956452537699824805835358067"), F("9.2076794298330791242156818474003")]],
    D=[_g23, F("0.24291996454816804366592249683314"), F("2.1851380027664058511513169485832")],
    B=[F(1), F("6.1697947043828245592553615689730"), F("-0.4277225654321857332623837380651")],
    E=[F("0.5"), F("-2.9079558716805469821718236208017"), F("0.2235406989781156962736090927619")])

TABLEAUS = {"OptRodas34": RODAS34, "OptRosenbrock34": ROS34, "OptRodas23": RODAS23, "OptRosenbrock23": ROS23}


def linear_step(tab, k, h, y0, with_d=True):
    """One attempt of the tableau on y' = J y, J = [[-k, 0], [k, 0]]:
    returns (y1, err) as mpmath vectors.  The reference clips every stage
    state at zero before evaluating f (FastChemistryModel.C:212-216), which
    this linear map does not do: the stage states must stay non-negative."""
    J = mp.matrix([[-k, 0], [k, 0]])
    G = mp.eye(2) / (tab["gamma"] * h) - J
    y0 = mp.matrix(y0)
    ks = []
    for s in range(len(tab["D"])):
        arg = y0 + sum((a * kj for a, kj in zip(tab["A"][s], ks)), mp.matrix([0, 0]))
        assert min(arg) > 0, "stage state below zero: the reference would clip it"
        rhs = J * arg + ((tab["D"][s] * h) if with_d else 0) * (J * y0)
        rhs = rhs + sum((c / h * kj for c, kj in zip(tab["C"][s], ks)), mp.matrix([0, 0]))
        ks.append(mp.lu_solve(G, rhs))
    y1 = y0 + sum((b * kj for b, kj in zip(tab["B"], ks)), mp.matrix([0, 0]))
    err = sum((e * kj for e, kj in zip(tab["E"], ks)), mp.matrix([0, 0]))
    return y1, err


@pytest.fixture(scope="module")
def linear_mech():
    """A -> B, irreversible, k = A_rate (beta = 0, Ta = 0), A and B with the
    same molar mass and NASA-7 fits (so the reaction releases no heat)."""
    def make(k):
        m = tiny_mech([R(M.RXN_ARRHENIUS, [0], [1], A=k, beta=0.0, Ta=0.0)], ns=2, seed=3)
        m.W = np.array([28.0, 28.0])
        m.coeffs_low[1] = m.coeffs_low[0]
        m.coeffs_high[1] = m.coeffs_high[0]
        m.T_common[1] = m.T_common[0]
        return m
    return make


ABS_TOL, REL_TOL = 1e-8, 1e-4


@pytest.mark.parametrize("solver", list(TABLEAUS))
@pytest.mark.parametrize("k,h", [(1e3, 1e-4), (1e4, 1e-4), (1e5, 1.5e-5), (1e2, 2e-3)])
def test_one_attempt_matches_linear_closed_form(linear_mech, solver, k, h):
    m = linear_mech(k)
    o = Oracle(m)
    y0 = np.array([0.625, 0.375, 1200.0])
    p = 1e5
    tab = TABLEAUS[solver]
    for textbook in (0, 1):
        orc_lib().orc_set_textbook_rosenbrock(textbook)
        try:
            norm, y1, _ = o.attempt(p, y0, h, solver=solver, absTol=ABS_TOL, relTol=REL_TOL)
        finally:
            orc_lib().orc_set_textbook_rosenbrock(0)
        ref, err = linear_step(tab, F(k), F(h), [F(0.625), F(0.375)], with_d=not textbook)
        other, _ = linear_step(tab, F(k), F(h), [F(0.625), F(0.375)], with_d=bool(textbook))
        for i in range(2):
            scale = abs(float(ref[i] - F(y0[i]))) + 1e-300     # the step's own increment
            assert abs(y1[i] - float(ref[i])) <= 1e-11 * scale, (solver, textbook, i, y1[i], float(ref[i]))
            # the d_i h f(y0) terms are visible at this bar: the other form misses it
            assert abs(float(other[i] - ref[i])) > 1e3 * 1e-11 * scale, (solver, textbook, i)
        assert y1[2] == y0[2]
        tol = [ABS_TOL + REL_TOL * max(abs(float(F(y0[i]))), abs(float(ref[i]))) for i in range(2)]
        norm_ref = max(abs(float(err[i])) / tol[i] for i in range(2))
        assert norm == pytest.approx(norm_ref, rel=1e-9, abs=1e-12), (solver, textbook, norm, norm_ref)
