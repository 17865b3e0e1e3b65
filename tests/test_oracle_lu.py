"""Pins the oracle's block-4 LU to the reference's own LUsolver.

tests/golden/lu_golden.npz was produced by the reference's
src/LUsolver/LUsolver.C compiled as-is (oracle/Makefile `ref`).
"""
import os

import numpy as np
import pytest

from oracle import lu_factor, lu_solve

GOLD = os.path.join(os.path.dirname(__file__), "golden", "lu_golden.npz")


def systems():
    g = np.load(GOLD)
    for i, n in enumerate(g["sizes"]):
        mo, vo = g["moff"], g["voff"]
        yield dict(i=i, n=int(n), kind=i % 4,
                   A=g["A"][mo[i]:mo[i + 1]].reshape(n, n), b=g["b"][vo[i]:vo[i + 1]],
                   x=g["x"][vo[i]:vo[i + 1]], piv=g["piv"][vo[i]:vo[i + 1]],
                   LU=g["LU"][mo[i]:mo[i + 1]].reshape(n, n))


@pytest.mark.parametrize("s", [s for s in systems() if s["kind"] != 3], ids=lambda s: f"n{s['n']}k{s['kind']}")
def test_lu_matches_reference(s):
    x, f, piv = lu_solve(s["A"], s["b"])
    np.testing.assert_array_equal(piv, s["piv"])
    np.testing.assert_allclose(f, s["LU"], rtol=0, atol=1e-11 * np.abs(s["LU"]).max())
    np.testing.assert_allclose(x, s["x"], rtol=0, atol=1e-10 * np.abs(s["x"]).max())


@pytest.mark.parametrize("s", [s for s in systems() if s["kind"] == 3], ids=lambda s: f"n{s['n']}")
def test_zero_pivot_replaced_like_reference(s):
    """A zero first column in a block: no swap, pivot replaced by 2^-52."""
    f, piv, invD = lu_factor(s["A"])
    assert piv[0] == s["piv"][0] == 0
    assert f[0, 0] == s["LU"][0, 0] == 2.2204460492503131e-16


def test_signed_running_max_quirk():
    """Column k0+1 candidates |1|, |-5|, |2|: the reference keeps the signed
    -5 as its running maximum and so picks the row holding 2."""
    a = np.eye(8)
    a[1, 1], a[2, 1], a[3, 1] = 1.0, -5.0, 2.0
    f, piv, _ = lu_factor(a)
    assert piv[1] == 3


@pytest.mark.parametrize("n", [5, 11, 54])
def test_lu_residual(n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + n * np.eye(n)
    b = rng.standard_normal(n)
    x, _, _ = lu_solve(A, b)
    np.testing.assert_allclose(A @ x, b, atol=1e-12 * n)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
