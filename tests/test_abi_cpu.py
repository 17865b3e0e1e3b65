"""CPU-side checks of the boundary: the C-ABI library loads and exports
every symbol include/fcs.h declares; the ctypes structures match the
header; host-side mechanism packing is consistent."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import fcs.mechanism as M
from fcs import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fcs.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fcs_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from fcs import lib
    L = lib.load()
    names = declared_functions()
    assert "fcs_solve" in names and len(names) >= 16
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(lib.EXPORTS)
    assert L.fcs_version().startswith(b"fcs")


def test_struct_layout_matches_header():
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct fcs_mech_desc \{(.*?)\} fcs_mech_desc;", src, re.S).group(1)
    fields = re.findall(r"(?:const )?(?:double|int32_t)\*?\s*\*?\s*(\w+)\s*;", body)
    assert fields == [f for f, _ in abi.MechDesc._fields_]
    body = re.search(r"typedef struct fcs_solver_opts \{(.*?)\} fcs_solver_opts;", src, re.S).group(1)
    fields = re.findall(r"(?:double|int32_t)\s+(\w+)\s*;", body)
    assert fields == [f for f, _ in abi.SolverOpts._fields_]


@pytest.mark.parametrize("cfg", [0, 1])
def test_mechanism_arrays_are_consistent(cfg):
    m = M.config_mechanism(cfg)
    a = m.arrays()
    ns, nr = a["n_species"], a["n_reactions"]
    assert a["rxn_lhs"].max() < ns and a["rxn_rhs"].max() < ns
    assert ((a["rxn_n_lhs"] >= 1) & (a["rxn_n_lhs"] <= 3)).all()
    nu = m.stoich_matrix()
    np.testing.assert_allclose(nu @ m.W, 0, atol=1e-10)          # mass balance
    np.testing.assert_array_equal(nu @ m.composition, 0)        # element balance
    assert (a["W"] >= 2.0).all() and (a["W"] <= 44.1).all()
    desc, keep = abi.make_desc(a)
    assert desc.n_species == ns and desc.n_reactions == nr


def test_configuration_sizes():
    m0, m1 = M.config_mechanism(0), M.config_mechanism(1)
    assert (m0.n_species, m0.n_reactions) == (10, 27)
    assert (m1.n_species, m1.n_reactions) == (53, 325)
    t1 = [r.rtype for r in m1.reactions]
    assert (t1.count(M.RXN_THIRD_BODY), t1.count(M.RXN_TROE), t1.count(M.RXN_PLOG)) == (60, 25, 5)
    assert [r.rtype for r in m0.reactions].count(M.RXN_THIRD_BODY) == 4
    f1 = M.config_field(1, m1, n_cells=4096)
    assert f1.T.min() >= 300.0 and f1.T.max() <= 2200.0
    np.testing.assert_allclose(f1.Y.sum(axis=0), 1.0, rtol=1e-12)


def test_generator_is_deterministic():
    a, b = M.config_mechanism(1).arrays(), M.config_mechanism(1).arrays()
    for k in a:
        np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]))


def test_bench_refuses_several_gpus():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
