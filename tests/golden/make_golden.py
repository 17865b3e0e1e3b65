"""Writes tests/golden/lu_golden.npz from the reference's own LUsolver.

The reference's src/LUsolver/LUsolver.C compiles as-is (oracle/Makefile,
target `ref`, output oracle/_ref/libref_lu.so).  This script runs it on
seeded matrices -- random ones and ones built to drive each branch of its
block-local pivot search -- and stores inputs, solutions, pivot indices and
the factored matrices as fixtures.  Run from the repository root:
    make -C oracle ref && python tests/golden/make_golden.py
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref_lu.so")


def ref_lu(a, b):
    lib = C.CDLL(LIB)
    lib.ref_lu_factor_solve.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_int,
                                        C.POINTER(C.c_double), C.POINTER(C.c_int)]
    n = a.shape[0]
    f = np.array(a, np.float64, order="C")
    x = np.array(b, np.float64, order="C")
    piv = np.zeros(n, np.int32)
    lib.ref_lu_factor_solve(n, f.ctypes.data_as(C.POINTER(C.c_double)), 1,
                            x.ctypes.data_as(C.POINTER(C.c_double)), piv.ctypes.data_as(C.POINTER(C.c_int)))
    return f, x, piv


def crafted(n, rng):
    """Matrices whose block columns exercise the pivot rules: a dominant
    first column (no swap), a swap in column k0, the signed running maximum
    in column k0+1 (|row1| < |row2|, row2 negative, |row3| < |row2|), a swap
    in column k0+2, and an exact zero pivot."""
    a = rng.standard_normal((n, n)) * 0.1
    for k0 in range(0, n - n % 4, 4):
        a[k0:k0 + 4, k0:k0 + 4] = 0.0
        a[k0, k0] = 1.0
        a[k0 + 1, k0 + 1] = 1.0
        a[k0 + 2, k0 + 1] = -5.0
        a[k0 + 3, k0 + 1] = 2.0
        a[k0 + 2, k0 + 2] = 0.5
        a[k0 + 3, k0 + 2] = 3.0
        a[k0 + 1, k0 + 3] = 1.0
        a[k0 + 2, k0 + 3] = 1.0
        a[k0 + 3, k0 + 3] = 1.0
    if n % 4:
        t = n - n % 4
        a[t:, t:] = rng.standard_normal((n - t, n - t))
    return a


def main():
    rng = np.random.default_rng(20240601)
    mats, rhs, xs, pivs, facs, sizes = [], [], [], [], [], []
    for n in (5, 8, 9, 10, 11, 12, 13, 14, 15, 54):
        for kind in range(4):
            if kind == 0:
                a = rng.standard_normal((n, n))
            elif kind == 1:     # I/(gamma h) - J like: dominant diagonal plus stiff entries
                a = np.diag(rng.uniform(1e3, 1e5, n)) - rng.standard_normal((n, n)) * 10 ** rng.uniform(0, 6, (n, n))
            elif kind == 2:
                a = crafted(n, rng)
            else:               # a block with an exact zero pivot candidate column
                a = rng.standard_normal((n, n))
                a[:4, 0] = 0.0
            b = rng.standard_normal(n)
            f, x, piv = ref_lu(a, b)
            mats.append(a.ravel()); rhs.append(b); xs.append(x); pivs.append(piv); facs.append(f.ravel())
            sizes.append(n)
    off = np.cumsum([0] + [s * s for s in sizes])
    offv = np.cumsum([0] + sizes)
    np.savez_compressed(os.path.join(HERE, "lu_golden.npz"), sizes=np.array(sizes, np.int32),
                        A=np.concatenate(mats), b=np.concatenate(rhs), x=np.concatenate(xs),
                        piv=np.concatenate(pivs).astype(np.int32), LU=np.concatenate(facs),
                        moff=off, voff=offv)
    print("wrote", len(sizes), "systems")


if __name__ == "__main__":
    main()
