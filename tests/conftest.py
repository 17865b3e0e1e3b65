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
