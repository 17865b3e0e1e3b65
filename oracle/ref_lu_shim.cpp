// ref_lu_shim.cpp -- extern "C" entry to the reference's own LUsolver
// (src/LUsolver/LUsolver.{H,C}, compiled as-is from the reference tree by
// oracle/Makefile into oracle/_ref/).  TEST INFRASTRUCTURE: it produces the
// golden vectors that pin the oracle's block-4 LU.
#include "LUsolver.H"

extern "C" int ref_lu_factor_solve(int n, double* a, int nrhs, double* b,
                                   int* piv)
{
    LUsolver lu(a, n);              // factors in place in a (row-major n x n)
    lu.Block4LUDecompose();
    for (int k = 0; k < nrhs; ++k) lu.xSolve(b + (long)k * n);
    for (int i = 0; i < n; ++i) piv[i] = (int)lu.pivotIndice(i);
    return 0;
}
