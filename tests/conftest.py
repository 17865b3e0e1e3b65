import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "fastchemistrysolver-openfoam-10_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def mech0():
    from fcs.mechanism import config_mechanism
    return config_mechanism(0)


@pytest.fixture(scope="session")
def mech1():
    from fcs.mechanism import config_mechanism
    return config_mechanism(1)


@pytest.fixture(scope="session")
def gri30():
    """GRI-3.0 (53 species, 325 reactions) from the reference's counterflow
    tutorial and 1,000 cells of its converged flame snapshot
    (tests/golden/gri30_counterflow.npz, written by tests/golden/make_gri30.py)."""
    import numpy as np
    from fcs.mechanism import Field, Mechanism
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "gri30_counterflow.npz"))
    a = {k[5:]: g[k] for k in g.files if k.startswith("mech_")}
    mech = Mechanism.from_arrays(a, names=[str(s) for s in g["names"]])
    field = Field(g["T"].copy(), g["p"].copy(), g["Y"].copy(), np.full(g["T"].size, 1e-7))
    return mech, field


TUTORIAL_CASES = ("NH3_Zhang_counterflow", "NC12H26_Yao_counterflow", "H2_Connaire2004",
                  "Kerosene_Zettervall2020", "NC12H26_Gao2014", "DME_LLNL", "NH3CH4_Konnov",
                  "NC7H16_Yoo2011", "NC12H26_Narayanaswamy2014", "C2H4_Sandiego", "C3H8_Qin2000")


@pytest.fixture(scope="session")
def tutorial_case():
    """Loader of the reference-tutorial mechanisms and cell states of
    tests/golden/tutorial_mechs.npz (written by tests/golden/
    make_tutorial_mechs.py): tutorial_case(name) -> (Mechanism, Field)."""
    import numpy as np
    from fcs.mechanism import Field, Mechanism
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "tutorial_mechs.npz"))
    cache = {}

    def load(name):
        if name not in cache:
            pre = name + "__mech_"
            a = {k[len(pre):]: g[k] for k in g.files if k.startswith(pre)}
            mech = Mechanism.from_arrays(a, names=[str(s) for s in g[name + "__names"]])
            T = g[name + "__T"]
            cache[name] = (mech, Field(T.copy(), g[name + "__p"].copy(), g[name + "__Y"].copy(),
                                       np.full(T.size, 1e-7)))
        m, f = cache[name]
        return m, Field(f.T.copy(), f.p.copy(), f.Y.copy(), f.deltaTChem.copy())

    return load
