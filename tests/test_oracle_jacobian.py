"""Pins the oracle's analytic Jacobians by central finite differences of the
restated right-hand sides, and the exact/fast dC/dY conversions by their
algebraic structure."""
import numpy as np
import pytest

import fcs.mechanism as M
from oracle import Oracle


def states(cfg, n):
    m = M.config_mechanism(cfg)
    fld = M.config_field(cfg, m, n_cells=max(n, 8))
    return m, [(fld.p[c], np.concatenate([fld.Y[:, c], [fld.T[c]]])) for c in range(n)]


def fd_jacobian(o, p, y, rel=1e-6):
    n = len(y)
    J = np.zeros((n, n))
    for j in range(n):
        h = rel * max(abs(y[j]), 1e-4) if j < n - 1 else 1e-3
        yp, ym = y.copy(), y.copy()
        yp[j] += h
        ym[j] -= h
        J[:, j] = (o.derivatives(p, yp)[0] - o.derivatives(p, ym)[0]) / (2 * h)
    return J


@pytest.mark.parametrize("cfg", [0, 1])
def test_exact_jacobian_matches_finite_differences(cfg):
    m, sts = states(cfg, 3)
    o = Oracle(m)
    for p, y in sts:
        y = np.maximum(y, 1e-3)   # keep every Y away from the clip at 0
        y[:-1] /= y[:-1].sum()
        f, J, _ = o.jacobian(p, y, "exact")
        Jfd = fd_jacobian(o, p, y)
        scale = np.abs(Jfd).max(axis=1, keepdims=True) + 1e-300
        err = np.abs(J - Jfd) / scale
        assert err.max() < 2e-5, err.max()
        np.testing.assert_allclose(f, o.derivatives(p, y)[0], rtol=1e-12, atol=1e-12 * np.abs(f).max())


@pytest.mark.parametrize("cfg", [0, 1])
def test_concentration_jacobian_matches_finite_differences(cfg):
    """Jc = d omega / d(C, T) against central differences of the rates."""
    m, sts = states(cfg, 2)
    o = Oracle(m)
    ns = m.n_species
    for p, y in sts:
        y = np.maximum(y, 1e-3)
        y[:-1] /= y[:-1].sum()
        f, J, yc, Jc = o.jacobian(p, y, "exact", want_jc=True)
        T = y[-1]
        rho = p / (M.RU * T * (yc[:-1] / m.W).sum())
        C = rho * yc[:-1] / m.W
        Jfd = np.zeros((ns, ns))
        for j in range(ns):
            h = 1e-6 * C[j]
            Cp, Cm = C.copy(), C.copy()
            Cp[j] += h
            Cm[j] -= h
            Jfd[:, j] = (o.rates(p, T, Cp) - o.rates(p, T, Cm)) / (2 * h)
        # row-scaled: a column entry far below its row's scale is lost to
        # cancellation inside q = kf CF - kr CR in the differences
        scale = np.abs(Jfd).max(axis=1, keepdims=True) + 1e-300
        assert (np.abs(Jfd - Jc[:ns, :ns]) / scale).max() < 1e-5
        hT = 1e-3
        colT = (o.rates(p, T + hT, C) - o.rates(p, T - hT, C)) / (2 * hT)
        assert np.abs(colT - Jc[:ns, ns]).max() / (np.abs(colT).max() + 1e-300) < 1e-5


@pytest.mark.parametrize("cfg", [0, 1])
def test_fast_differs_from_exact_by_rank_one(cfg):
    """exact - fast = -(W_i/rho)(MW/W_j) sum_k Jc_ik C_k on the species block:
    a rank-1 matrix (diagonal dC/dY drops exactly that term)."""
    m, sts = states(cfg, 2)
    o = Oracle(m)
    ns = m.n_species
    for p, y in sts:
        fe, Je, _ = o.jacobian(p, y, "exact")
        ff, Jf, _ = o.jacobian(p, y, "fast")
        np.testing.assert_array_equal(fe, ff)
        D = (Je - Jf)[:ns, :ns]
        s = np.linalg.svd(D, compute_uv=False)
        assert s[1] <= 1e-9 * s[0]
        # the T column is the same in both forms
        np.testing.assert_allclose(Je[:ns, ns], Jf[:ns, ns], rtol=1e-13, atol=0)


def test_rodas_matrix_is_well_posed(mech1):
    """I/(gamma h) - J is solved by the oracle LU to a small residual."""
    from oracle import lu_solve
    m, sts = states(1, 2)
    o = Oracle(m)
    for p, y in sts:
        f, J, _ = o.jacobian(p, y)
        for h in (1e-9, 1e-7, 1e-6):
            A = np.eye(len(y)) / (0.25 * h) - J
            x, _, _ = lu_solve(A, f)
            r = A @ x - f
            assert np.abs(r).max() <= 1e-9 * (np.abs(A).max() * np.abs(x).max() + np.abs(f).max())
