"""The library's mechanism-table build (csrc/fcs_mech_tables.h, the code
fcs_mech_create runs, built for the host in tests/emu) on every mechanism
the GPU tests use, at the padded size the library picks (both forms): it
must accept them all, its layout checks included (slot counts and species
indices, PLOG rows in sorted order, reaction-word fields in range, the
per-species list bounds)."""
import numpy as np
import pytest

import fcs.mechanism as M
from conftest import TUTORIAL_CASES

from emu import emu

SIZES = (12, 24, 40, 56, 64, 80, 96, 112, 128, 176, 256)


def padded(ns):
    return next(n for n in SIZES if n >= ns + 1)


@pytest.mark.parametrize("cfg", [0, 1, 4])
def test_config_mechanisms(cfg):
    mech = M.config_mechanism(cfg)
    assert emu.tables(mech, padded(mech.n_species)) == 0


def test_gri30(gri30):
    mech, _ = gri30
    assert emu.tables(mech, padded(mech.n_species)) == 0


def test_tutorial_mechanisms(tutorial_case):
    for name in TUTORIAL_CASES:
        mech, _ = tutorial_case(name)
        assert emu.tables(mech, padded(mech.n_species)) == 0, name
