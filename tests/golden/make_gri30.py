"""Writes tests/golden/gri30_counterflow.npz from the reference's own
tutorial data (run where the reference tree is present):

* the GRI-3.0 mechanism (53 species, 325 reactions) of
  tutorial/CounterFlowFlame2D/CH4_GRI30/constant/{reactions,speciesThermo},
  parsed by fcs.foam and stored as the flat fcs_mech_desc arrays;
* every 4th cell (1,000 of 4,000) of the converged counterflow flame
  snapshot tutorial/CounterFlowFlame2D/CH4_GRI30/1/ (T, p, and the 53 mass
  fractions).

The fixture is data only: mechanism coefficients and field values.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "fastchemistrysolver-openfoam-10_amd"))
from fcs import foam  # noqa: E402

REF = os.environ.get("FCS_REFERENCE", "/root/reference")
CASE = os.path.join(REF, "tutorial", "CounterFlowFlame2D", "CH4_GRI30")


def main():
    mech = foam.load_mechanism(os.path.join(CASE, "constant", "reactions"),
                               os.path.join(CASE, "constant", "speciesThermo"))
    a = mech.arrays()
    T = foam.read_field(os.path.join(CASE, "1", "T"))
    n = T.size
    p = foam.read_field(os.path.join(CASE, "1", "p"), n)
    Y = np.array([foam.read_field(os.path.join(CASE, "1", s), n) for s in mech.names])
    pick = np.arange(0, n, 4)
    out = {f"mech_{k}": np.asarray(v) for k, v in a.items()}
    out["names"] = np.array(mech.names)
    out["T"], out["p"], out["Y"] = T[pick], p[pick], Y[:, pick]
    np.savez_compressed(os.path.join(HERE, "gri30_counterflow.npz"), **out)
    print(f"{mech.n_species} species, {mech.n_reactions} reactions, {pick.size} cells")


if __name__ == "__main__":
    main()
