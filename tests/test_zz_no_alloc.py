"""fcs_solve_noalloc, the no-allocation variant of the batch solve
(fcs_batch_upload + fcs_batch_run + fcs_batch_download over device arrays
the caller holds), against fcs_batch_run on the same cells: bit-identical
results on two calls in a row, the second with fewer cells (the scratch keeps
the first call's contents); inputs beyond the scratch are refused.  The test
holds its device arrays through the HIP runtime libfcs.so itself links
(hipMalloc / hipMemcpy / hipFree over ctypes), as a caller would."""
import ctypes as C

import numpy as np
import pytest

import fcs.mechanism as M

pytestmark = pytest.mark.gpu

DT = 1e-6


class DeviceArrays:
    """Device copies of host arrays, through the library's own HIP runtime."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.held = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.held.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0     # host to device
        return p.value

    def get(self, ptr, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.held:
            self.hip.hipFree(p)
        self.held = []


def test_noalloc_matches_batch_on_two_calls():
    from fcs import lib
    from fcs.abi import FCS_E_SCRATCH, FCS_OK
    L = lib.load()
    assert lib.device_count() >= 1
    dev = DeviceArrays()
    mech = M.config_mechanism(1)
    s = lib.Solver(lib.Mech(mech), Treact=500.0)
    full = M.config_field(1, mech, n_cells=4096)
    full.T[::7] = 450.0                      # cold cells, skipped and written out
    rho_full = full.rho(mech)
    try:
        for n in (4096, 1000):
            f = full.subset(np.arange(n))
            print(f"noalloc: {n} cells, {mech.n_species} species", flush=True)
            b = lib.Batch(s, f)
            b.run(DT)
            ref = b.download()
            T0, p0, rho0, Y0, dtc = (dev.put(a) for a in (f.T, f.p, rho_full[:n], f.Y, f.deltaTChem))
            dtcOut = dev.put(np.full(n, -1.0))
            RR = dev.put(np.full((mech.n_species, n), np.nan))
            st = dev.put(np.full(n, -1, np.int32))
            Tend = dev.put(np.zeros(n))
            dmin = np.zeros(1)
            rc = L.fcs_solve_noalloc(s._h, n, DT, T0, p0, rho0, Y0, dtc, dtcOut, RR, st, Tend, lib.dptr(dmin))
            assert rc == FCS_OK
            np.testing.assert_array_equal(dev.get(RR, ref["RR"]), ref["RR"])
            np.testing.assert_array_equal(dev.get(dtcOut, ref["deltaTChem"]), ref["deltaTChem"])
            np.testing.assert_array_equal(dev.get(st, ref["status"]), ref["status"])
            np.testing.assert_array_equal(dev.get(Tend, ref["Tend"]), ref["Tend"])
            assert dmin[0] == ref["deltaTMin"]
            assert (ref["status"][f.T <= 500.0] == 1).all() and (ref["status"][f.T > 500.0] == 0).all()
        # beyond the 32 MiB scratch: refused before any launch
        rc = L.fcs_solve_noalloc(s._h, 10 ** 6, DT, T0, p0, rho0, Y0, dtc, dtcOut, RR, st, Tend, None)
        assert rc == FCS_E_SCRATCH
    finally:
        dev.free()
