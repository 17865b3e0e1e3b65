"""Algorithmic FP64 work of one Rosenbrock attempt on one cell.

Counting rules: add/mul = 1, FMA = 2, div/sqrt = 1, each exp/log/pow = 1.
The conversion dC/dY is counted in its rank-1 form (the algorithmic
minimum), so redundant O(n^3) work earns no credit.  The figure is a
property of the mechanism: sizes and reaction forms.
"""
from __future__ import annotations

from .mechanism import RXN_ARRHENIUS, RXN_LINDEMANN, RXN_PLOG, RXN_THIRD_BODY, RXN_TROE


def _rxn_counts(mech):
    c = {t: 0 for t in (RXN_ARRHENIUS, RXN_THIRD_BODY, RXN_LINDEMANN, RXN_TROE, RXN_PLOG)}
    for r in mech.reactions:
        c[r.rtype] += 1
    return c


def rhs_flops(mech) -> int:
    """One right-hand side evaluation (FastChemistryModel::derivatives)."""
    ns, nr = mech.n_species, mech.n_reactions
    c = _rxn_counts(mech)
    n_tb = c[RXN_THIRD_BODY] + c[RXN_LINDEMANN] + c[RXN_TROE]
    f = 0
    f += 3 * ns + 2          # clip, density sum, 1/rho
    f += ns * 2              # concentrations
    f += ns * 20             # cp, h, -G/RT polynomials + exp
    f += nr * 6              # kf = A exp(beta lnT - Ta/T)
    f += n_tb * 2 * ns       # third-body sums
    f += c[RXN_LINDEMANN] * 6
    f += c[RXN_TROE] * 40    # Fcent (3 exp), 2 log10, pow, rational terms
    f += c[RXN_PLOG] * 16
    f += nr * 12             # Kc products, floor, kr, CF, CR, q
    f += nr * 2 * 4          # scatter/gather of q into the species
    f += ns * 6              # Ydot, cpM, dT/dt
    return f


def jacobian_flops(mech) -> int:
    """Jacobian (Jc, rank-1 dC/dY conversion, T row/column) plus f(y0)."""
    ns, nr = mech.n_species, mech.n_reactions
    n = ns + 1
    c = _rxn_counts(mech)
    n_tb = c[RXN_THIRD_BODY] + c[RXN_LINDEMANN] + c[RXN_TROE]
    f = rhs_flops(mech)
    f += nr * 8              # dkf/dT, dkr/dT, dq/dT
    f += c[RXN_TROE] * 45    # dF/dT, dF/dPr
    f += c[RXN_LINDEMANN] * 10
    f += nr * 4 * 6          # sparse partials of q into the species rows
    f += n_tb * 4 * 2 * ns   # third-body partial rows (about 4 species per reaction)
    f += ns * (2 * ns)       # s_i = sum_k Jc_ik C_k
    f += ns * ns * 6         # rank-1 conversion
    f += ns * 6              # T column
    f += n * 2 * ns + 2 * ns  # T row
    return f


def lu_flops(n: int) -> int:
    return sum((n - k - 1) + 2 * (n - k - 1) ** 2 for k in range(n))


def trsv_flops(n: int) -> int:
    return 2 * n * n - n


# (triangular solves, RHS evaluations, vector updates per row) of one attempt
_STAGES = {"OptRodas34": (6, 5, 12), "OptRosenbrock34": (4, 2, 10),
           "OptRodas23": (4, 2, 10), "OptRosenbrock23": (3, 1, 8)}


def attempt_flops(mech, solver: str = "OptRodas34") -> int:
    """FP64 operations of one attempt: one Jacobian and one LU, then
    Rodas34 6 solves / 5 RHS, Rosenbrock34 4 / 2, Rodas23 4 / 2,
    Rosenbrock23 3 / 1."""
    n = mech.n_species + 1
    solves, rhs, vec = _STAGES[solver]
    return jacobian_flops(mech) + lu_flops(n) + solves * trsv_flops(n) + rhs * rhs_flops(mech) + vec * n


def hbm_bytes_per_cell(mech) -> int:
    """Algorithmic HBM traffic of one cell and one flow step: read T0, p0,
    rho0, deltaTChem and Y0[ns]; write RR[ns], deltaTChem, status, T."""
    ns = mech.n_species
    return 8 * (ns + 4) + 8 * (ns + 2) + 4
