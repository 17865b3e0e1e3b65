"""Host build of the library's wavefront-form step kernel (emu_wave.cpp,
include/hip/hip_runtime.h) and its ctypes binding, for the CPU tests: the
device source runs with one thread per lane of a 64-lane wavefront, every
cross-lane operation a rendezvous of the 64 lanes that checks that they all
came from the same call site with the same lane argument."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(REPO, "fastchemistrysolver-openfoam-10_amd", "csrc")
OUT = os.path.join(HERE, "_build", "libfcs_emu.so")
CXX = os.environ.get("FCS_EMU_CXX", "/opt/rocm/llvm/bin/clang++")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def build():
    """Compile emu_wave.cpp (seconds to a minute); returns the library path."""
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    srcs = [os.path.join(HERE, "emu_wave.cpp"), os.path.join(HERE, "include", "hip", "hip_runtime.h")]
    srcs += [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    if os.path.exists(OUT) and os.path.getmtime(OUT) > max(os.path.getmtime(s) for s in srcs):
        return OUT
    tmp = f"{OUT}.{os.getpid()}.tmp"   # concurrent test workers each build their own copy
    cmd = [CXX, "-O1", "-std=c++17", "-shared", "-fPIC", "-pthread", "-Wno-pass-failed",
           "-I" + os.path.join(HERE, "include"), "-I" + os.path.join(REPO, "include"), "-I" + CSRC,
           os.path.join(HERE, "emu_wave.cpp"), "-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("emulator build failed:\n" + r.stdout[-4000:])
    os.replace(tmp, OUT)
    return OUT


def solve(mech, field, deltaT, solver="OptRodas34", absTol=1e-8, relTol=1e-4, maxSteps=10000, Treact=0.0,
          dtChemMax=1.0e15, timeout=60.0, P=None, seed=0):
    """solve(deltaT) of every cell of the field through the host build of
    k_step (k_step_seulex for OptSeulex); returns the results and the
    rendezvous counts.  P (float32 [reacting cells, ns+2], rows in cell
    order: [ln h, z_0 .. z_ns-1, z_T]) runs the learned seed first: the
    gate, then k_verify's Rodas34 attempt per gated proposal (seed != 0:
    rejected and ungated cells start from min(h, deltaT)); nAccepted is
    [accepted, gated]."""
    from fcs.abi import SOLVER_NAMES, make_desc
    L = C.CDLL(build())
    L.emu_solve.restype = C.c_int
    desc, keep = make_desc(mech.arrays())
    nc, ns = field.n_cells, mech.n_species
    T0 = np.ascontiguousarray(field.T, np.float64)
    p0 = np.ascontiguousarray(field.p, np.float64)
    rho0 = np.ascontiguousarray(field.rho(mech), np.float64)
    Y0 = np.ascontiguousarray(field.Y, np.float64)
    dtc = np.array(field.deltaTChem, np.float64)
    RR = np.zeros((ns, nc))
    Tend = np.zeros(nc)
    dto = np.zeros(nc)
    st = np.zeros(nc, np.int32)
    na = np.zeros(nc, np.int32)
    checks, viol = C.c_longlong(0), C.c_longlong(0)
    msg = C.create_string_buffer(512)
    d = lambda a: a.ctypes.data_as(_dp)
    i = lambda a: a.ctypes.data_as(_ip)
    if P is not None:
        P = np.ascontiguousarray(P, np.float32)
        assert P.ndim == 2 and P.shape[1] == ns + 2 and P.shape[0] == int((field.T > Treact).sum())
    nacc = (C.c_ulonglong * 2)()
    rc = L.emu_solve(C.byref(desc), C.c_int(SOLVER_NAMES[solver]), C.c_double(absTol), C.c_double(relTol), C.c_int(maxSteps), C.c_double(Treact),
                     C.c_double(dtChemMax), C.c_int(nc), C.c_double(deltaT), d(T0), d(p0), d(rho0), d(Y0), d(dtc),
                     d(RR), d(Tend), d(dto), i(st), i(na), C.byref(checks), C.byref(viol), msg, C.c_int(512),
                     C.c_double(timeout), None if P is None else P.ctypes.data_as(C.POINTER(C.c_float)),
                     C.c_int(seed), nacc)
    return dict(rc=rc, RR=RR, Tend=Tend, deltaTChem=dto, status=st, nAttempts=na, checks=checks.value,
                violations=viol.value, message=msg.value.decode(), nAccepted=(nacc[0], nacc[1]))


def derivatives(mech, p, YT):
    """RHS of the states YT [ns+1, nS] at pressures p [nS] through the host
    build of k_derivatives; returns (dYT, YT clipped, checks, violations, message)."""
    from fcs.abi import make_desc
    L = C.CDLL(build())
    desc, keep = make_desc(mech.arrays())
    YT = np.array(YT, np.float64, order="C")
    p = np.ascontiguousarray(p, np.float64)
    out = np.zeros_like(YT)
    checks, viol = C.c_longlong(0), C.c_longlong(0)
    msg = C.create_string_buffer(512)
    rc = L.emu_derivatives(C.byref(desc), C.c_int(YT.shape[1]), p.ctypes.data_as(_dp), YT.ctypes.data_as(_dp),
                           out.ctypes.data_as(_dp), C.byref(checks), C.byref(viol), msg, C.c_int(512))
    return dict(rc=rc, d=out, YT=YT, checks=checks.value, violations=viol.value, message=msg.value.decode())


def jacobian(mech, p, YT, jac="exact"):
    """Jacobian of the states YT [ns+1, nS] through the host build of
    k_jacobian; returns (f [n, nS], J [nS, n, n], checks, violations, message)."""
    from fcs.abi import make_desc
    L = C.CDLL(build())
    desc, keep = make_desc(mech.arrays())
    YT = np.array(YT, np.float64, order="C")
    n, nS = YT.shape
    p = np.ascontiguousarray(p, np.float64)
    f = np.zeros_like(YT)
    J = np.zeros((nS, n, n))
    checks, viol = C.c_longlong(0), C.c_longlong(0)
    msg = C.create_string_buffer(512)
    rc = L.emu_jacobian(C.byref(desc), C.c_int(0 if jac == "exact" else 1), C.c_int(nS), p.ctypes.data_as(_dp),
                        YT.ctypes.data_as(_dp), f.ctypes.data_as(_dp), J.ctypes.data_as(_dp), C.byref(checks),
                        C.byref(viol), msg, C.c_int(512))
    return dict(rc=rc, f=f, J=J, YT=YT, checks=checks.value, violations=viol.value, message=msg.value.decode())


def attempt(mech, p, YT, dx, absTol=1e-8, relTol=1e-4):
    """One Rodas34 attempt of step dx from each state through the host build of
    k_attempt; returns (YTnew, err, checks, violations, message)."""
    from fcs.abi import make_desc
    L = C.CDLL(build())
    desc, keep = make_desc(mech.arrays())
    YT = np.array(YT, np.float64, order="C")
    n, nS = YT.shape
    p = np.ascontiguousarray(p, np.float64)
    dx = np.ascontiguousarray(np.broadcast_to(dx, (nS,)), np.float64)
    YTn = np.zeros_like(YT)
    err = np.zeros(nS)
    checks, viol = C.c_longlong(0), C.c_longlong(0)
    msg = C.create_string_buffer(512)
    rc = L.emu_attempt(C.byref(desc), C.c_double(absTol), C.c_double(relTol), C.c_int(nS), p.ctypes.data_as(_dp),
                       dx.ctypes.data_as(_dp), YT.ctypes.data_as(_dp), YTn.ctypes.data_as(_dp),
                       err.ctypes.data_as(_dp), C.byref(checks), C.byref(viol), msg, C.c_int(512))
    return dict(rc=rc, YTnew=YTn, err=err, YT=YT, checks=checks.value, violations=viol.value,
                message=msg.value.decode())


def selftest(kind, timeout=5.0):
    """Runs one of the checker's own cases (see emu_selftest)."""
    L = C.CDLL(build())
    checks, viol = C.c_longlong(0), C.c_longlong(0)
    msg = C.create_string_buffer(512)
    rc = L.emu_selftest(C.c_int(kind), C.c_double(timeout), C.byref(checks), C.byref(viol), msg, C.c_int(512))
    return dict(rc=rc, checks=checks.value, violations=viol.value, message=msg.value.decode())


def tables(mech, NP):
    """Return code of the library's table build (fcs_mech_tables.h) for the
    padded size NP."""
    from fcs.abi import make_desc
    L = C.CDLL(build())
    desc, keep = make_desc(mech.arrays())
    return L.emu_tables(C.byref(desc), C.c_int(NP))
