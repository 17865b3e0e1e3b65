"""Writes tests/golden/tutorial_mechs.npz from the reference's own tutorial
data (run where the reference tree is present).

For each case below: the mechanism of tutorial/<case>/constant/{reactions,
speciesThermo}, parsed by fcs.foam and stored as the flat fcs_mech_desc
arrays, and a set of cell states:

* counterflow cases: every k-th cell (256 cells) of the flame snapshot in
  tutorial/<case>/1/ (T, p and the mass fractions as the tutorial holds them);
* ignition cases: the mixture of constant/initialConditions (mole or mass
  fractions, p) at 64 temperatures drawn uniformly in [900, 2500] K, each with
  a radical pool of 1e-3 mass fraction spread over all species (Dirichlet
  weights, seeded), so that every reaction of the mechanism runs.

The cases cover the reaction forms and layouts the tutorials use:
ArrheniusPLOG (NH3_Zhang), nonEquilibrium reversible and third-body rows
(H2_Connaire2004, DME_LLNL), irreversible reactions with 4 to 7 product
slots (NC12H26_Yao, Kerosene_Zettervall2020, NC12H26_Gao2014), SRI falloff
(NH3CH4_Konnov), a Kc-reversible reaction with 4 product slots
(C3H8_Qin2000) and species counts across the padded system sizes of both
execution forms (10 to 255).

The fixture is data only: mechanism coefficients and field values.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "fastchemistrysolver-openfoam-10_amd"))
from fcs import foam  # noqa: E402

REF = os.environ.get("FCS_REFERENCE", "/root/reference")
CASES = {
    "NH3_Zhang_counterflow": "CounterFlowFlame2D/NH3_Zhang",
    "NC12H26_Yao_counterflow": "CounterFlowFlame2D/NC12H26_Yao",
    "H2_Connaire2004": "Ignition0D/H2_Connaire2004",
    "Kerosene_Zettervall2020": "Ignition0D/Kerosene_Zettervall2020",
    "NC12H26_Gao2014": "Ignition0D/NC12H26_Gao2014",
    "DME_LLNL": "Ignition0D/DME_LLNL",
    "NH3CH4_Konnov": "Ignition0D/NH3CH4_Konnov",
    "NC7H16_Yoo2011": "Ignition0D/NC7H16_Yoo2011",
    "NC12H26_Narayanaswamy2014": "Ignition0D/NC12H26_Narayanaswamy2014",
    "C2H4_Sandiego": "Ignition0D/C2H4_Sandiego",
    "C3H8_Qin2000": "Ignition0D/C3H8_Qin2000",
}
N_SNAPSHOT, N_IGNITION = 256, 64


def snapshot_states(case, mech):
    T = foam.read_field(os.path.join(case, "1", "T"))
    n = T.size
    p = foam.read_field(os.path.join(case, "1", "p"), n)
    Y = []
    for s in mech.names:
        f = os.path.join(case, "1", s)
        Y.append(foam.read_field(f, n) if os.path.exists(f) else np.zeros(n))
    pick = np.linspace(0, n - 1, N_SNAPSHOT).round().astype(int)
    return T[pick], p[pick], np.array(Y)[:, pick]


def ignition_states(case, mech, seed):
    ic = foam.read(os.path.join(case, "constant", "initialConditions"))
    x = np.zeros(mech.n_species)
    for name, v in ic["fractions"].items():
        x[mech.names.index(name)] = float(v)
    x /= x.sum()
    y = x * mech.W / (x * mech.W).sum() if ic.get("fractionBasis", "mole") == "mole" else x
    rng = np.random.default_rng(seed)
    T = rng.uniform(900.0, 2500.0, N_IGNITION)
    pool = rng.dirichlet(np.ones(mech.n_species), N_IGNITION).T * 1e-3
    Y = (1.0 - 1e-3) * y[:, None] + pool
    p = np.full(N_IGNITION, float(ic["p"]))
    return T, p, Y


def main():
    out = {}
    for i, (key, rel) in enumerate(CASES.items()):
        case = os.path.join(REF, "tutorial", rel)
        mech = foam.load_mechanism(os.path.join(case, "constant", "reactions"),
                                   os.path.join(case, "constant", "speciesThermo"))
        if os.path.isdir(os.path.join(case, "1")):
            T, p, Y = snapshot_states(case, mech)
        else:
            T, p, Y = ignition_states(case, mech, seed=1000 + i)
        for k, v in mech.arrays().items():
            out[f"{key}__mech_{k}"] = np.asarray(v)
        out[f"{key}__names"] = np.array(mech.names)
        out[f"{key}__T"], out[f"{key}__p"], out[f"{key}__Y"] = T, p, Y
        print(f"{key}: {mech.n_species} species, {mech.n_reactions} reactions, {T.size} cells")
    np.savez_compressed(os.path.join(HERE, "tutorial_mechs.npz"), **out)


if __name__ == "__main__":
    main()
