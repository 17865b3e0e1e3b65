"""ctypes mirror of the structures declared in include/fcs.h."""
from __future__ import annotations

import ctypes as C

import numpy as np

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

FCS_OK, FCS_E_ARG, FCS_E_HIP, FCS_E_UNSUPPORTED, FCS_E_NOMEM, FCS_E_CELLS, FCS_E_SCRATCH = 0, -1, -2, -3, -4, -5, -6
FCS_CELL_OK, FCS_CELL_SKIPPED, FCS_CELL_MAXSTEPS, FCS_CELL_NONFINITE = 0, 1, 2, 4
FCS_SOLVER_RODAS34, FCS_SOLVER_ROSENBROCK34, FCS_SOLVER_SEULEX = 0, 1, 2
FCS_SOLVER_RODAS23, FCS_SOLVER_ROSENBROCK23 = 3, 4
FCS_JACOBIAN_EXACT, FCS_JACOBIAN_FAST = 0, 1

SOLVER_NAMES = {"OptRodas34": FCS_SOLVER_RODAS34, "OptRosenbrock34": FCS_SOLVER_ROSENBROCK34,
                "OptSeulex": FCS_SOLVER_SEULEX, "OptRodas23": FCS_SOLVER_RODAS23,
                "OptRosenbrock23": FCS_SOLVER_ROSENBROCK23}
JACOBIAN_NAMES = {"exact": FCS_JACOBIAN_EXACT, "fast": FCS_JACOBIAN_FAST}


class MechDesc(C.Structure):
    _fields_ = [
        ("n_species", C.c_int32), ("n_reactions", C.c_int32),
        ("W", _dp), ("T_low", _dp), ("T_high", _dp), ("T_common", _dp),
        ("coeffs_high", _dp), ("coeffs_low", _dp),
        ("rxn_type", _ip), ("rxn_reversible", _ip), ("rxn_n_lhs", _ip), ("rxn_n_rhs", _ip),
        ("rxn_lhs", _ip), ("rxn_rhs", _ip),
        ("A", _dp), ("beta", _dp), ("Ta", _dp), ("A0", _dp), ("beta0", _dp), ("Ta0", _dp),
        ("efficiencies", _dp), ("troe", _dp),
        ("plog_max_points", C.c_int32), ("plog_n", _ip), ("plog_p", _dp),
        ("plog_A", _dp), ("plog_beta", _dp), ("plog_Ta", _dp),
        ("sri", _dp), ("A_rev", _dp), ("beta_rev", _dp), ("Ta_rev", _dp),
        ("slot_stride", C.c_int32),
    ]


class SolverOpts(C.Structure):
    _fields_ = [("solver", C.c_int32), ("jacobian", C.c_int32), ("absTol", C.c_double),
                ("relTol", C.c_double), ("maxSteps", C.c_int32), ("Treact", C.c_double),
                ("deltaTChemMax", C.c_double)]


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def make_desc(arrays: dict):
    """Build an MechDesc over the arrays of ``Mechanism.arrays()``.  Returns
    (desc, keepalive) -- keep the second value alive while desc is used."""
    keep = {}
    d = MechDesc()
    for name, typ in MechDesc._fields_:
        if name == "slot_stride" and name not in arrays:
            continue                     # 0: three slots per reaction
        v = arrays[name]
        if typ in (C.c_int32,):
            setattr(d, name, int(v))
            continue
        if typ is _dp:
            a = np.ascontiguousarray(v, np.float64)
            setattr(d, name, dptr(a))
        else:
            a = np.ascontiguousarray(v, np.int32)
            setattr(d, name, iptr(a))
        keep[name] = a
    return d, keep


def make_opts(solver="OptRodas34", jacobian="exact", absTol=1e-8, relTol=1e-4,
              maxSteps=10000, Treact=0.0, deltaTChemMax=1.0e15) -> SolverOpts:
    s = SOLVER_NAMES[solver] if isinstance(solver, str) else int(solver)
    j = JACOBIAN_NAMES[jacobian] if isinstance(jacobian, str) else int(jacobian)
    return SolverOpts(s, j, absTol, relTol, maxSteps, Treact, deltaTChemMax)


def row_pointers(a2d: np.ndarray):
    """Array of per-row double* for a C-contiguous [rows, n] array."""
    assert a2d.flags.c_contiguous and a2d.dtype == np.float64
    rows = a2d.shape[0]
    base = a2d.ctypes.data
    stride = a2d.strides[0]
    arr = (_dp * rows)()
    for i in range(rows):
        arr[i] = C.cast(base + i * stride, _dp)
    return arr
