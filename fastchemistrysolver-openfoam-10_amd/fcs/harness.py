"""Case files and driver of the adapter harness (csrc/fcs_harness.cpp), the
host program that stands in for OpenFOAM around include/fcs_chemistry_model.hpp.

A case directory holds options.txt ("key value" lines; sub-dictionary keys
as sub.key, e.g. OptRodas34Coeffs.absTol), mech.bin (the fcs_mech_desc
arrays) and fields.bin (cells, T, p, rho, Y per species)."""
from __future__ import annotations

import os
import subprocess

import numpy as np

from .build import HARNESS
from .mechanism import Field, Mechanism

_DOUBLE_HEAD = ("W", "T_low", "T_high", "T_common", "coeffs_high", "coeffs_low")
_INT_MID = ("rxn_type", "rxn_reversible", "rxn_n_lhs", "rxn_n_rhs", "rxn_lhs", "rxn_rhs")
_DOUBLE_MID = ("A", "beta", "Ta", "A0", "beta0", "Ta0", "efficiencies", "troe")
_DOUBLE_TAIL = ("plog_p", "plog_A", "plog_beta", "plog_Ta", "sri", "A_rev", "beta_rev", "Ta_rev")


def flatten(d: dict, prefix: str = "") -> dict:
    """A parsed OpenFOAM dictionary (fcs.foam.parse) as flat "sub.key" -> text."""
    out = {}
    for k, v in d.items():
        if isinstance(v, dict):
            out.update(flatten(v, prefix + k + "."))
        elif not isinstance(v, list):
            out[prefix + k] = repr(v) if isinstance(v, float) else str(v)
    return out


def write_case(path: str, mech: Mechanism, field: Field, options: dict, rho=None) -> None:
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "options.txt"), "w") as f:
        for k, v in options.items():
            f.write(f"{k} {v}\n")
    a = mech.arrays()
    with open(os.path.join(path, "mech.bin"), "wb") as f:
        f.write(np.array([a["n_species"], a["n_reactions"], a["plog_max_points"], a["slot_stride"]],
                         np.int32).tobytes())
        for k in _DOUBLE_HEAD:
            f.write(np.ascontiguousarray(a[k], np.float64).tobytes())
        for k in _INT_MID:
            f.write(np.ascontiguousarray(a[k], np.int32).tobytes())
        for k in _DOUBLE_MID:
            f.write(np.ascontiguousarray(a[k], np.float64).tobytes())
        f.write(np.ascontiguousarray(a["plog_n"], np.int32).tobytes())
        for k in _DOUBLE_TAIL:
            f.write(np.ascontiguousarray(a[k], np.float64).tobytes())
    r = field.rho(mech) if rho is None else np.asarray(rho, np.float64)
    with open(os.path.join(path, "fields.bin"), "wb") as f:
        f.write(np.array([field.n_cells], np.int64).tobytes())
        for v in (field.T, field.p, r, field.Y):
            f.write(np.ascontiguousarray(v, np.float64).tobytes())


def options(path: str) -> dict:
    """The options the adapter reads from the case (no device use)."""
    r = subprocess.run([HARNESS, path, "--options-only"], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr.strip())
    out = {}
    for line in r.stdout.splitlines():
        k, v = line.split("=", 1)
        out[k] = v if k == "solver" else float(v)
    return out


def run(path: str, deltaT: float, n_steps: int, ns: int, n_cells: int, lts: bool = False) -> dict:
    """Runs the harness's time loop; returns its outputs."""
    out = os.path.join(path, "out.bin")
    cmd = [HARNESS, path, repr(float(deltaT)), str(int(n_steps)), out] + (["--lts"] if lts else [])
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"harness failed ({r.returncode}): {r.stderr.strip()}")
    v = np.fromfile(out, np.float64)
    o = {}
    i = 0
    for k, n in (("deltaTMin", n_steps), ("Y", ns * n_cells), ("deltaTChem", n_cells), ("RR", ns * n_cells),
                 ("RR_calculate", ns * n_cells), ("tc", n_cells), ("Qdot", n_cells)):
        o[k] = v[i:i + n]
        i += n
    for k in ("Y", "RR", "RR_calculate"):
        o[k] = o[k].reshape(ns, n_cells)
    return o
