"""ctypes binding of libfcs.so (include/fcs.h).

The library is the only compute path: if it is missing or no GPU is
visible, the calls raise -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .abi import (FCS_E_CELLS, FCS_OK, MechDesc, SolverOpts, dptr, iptr, make_desc, make_opts,
                  row_pointers)
from .build import LIB

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_vp = C.c_void_p

EXPORTS = [
    "fcs_device_count", "fcs_version", "fcs_mech_create", "fcs_mech_destroy", "fcs_solver_create",
    "fcs_solver_destroy", "fcs_solve", "fcs_solve_ex", "fcs_solve_lts", "fcs_batch_create", "fcs_batch_destroy",
    "fcs_batch_upload", "fcs_batch_run", "fcs_batch_profile", "fcs_batch_stats", "fcs_batch_load_report",
    "fcs_batch_download",
    "fcs_eval_derivatives", "fcs_eval_jacobian", "fcs_eval_attempt", "fcs_lu_solve",
    "fcs_batch_set_learned", "fcs_batch_learned_stats", "fcs_batch_first_steps", "fcs_batch_set_launch",
    "fcs_batch_autotune", "fcs_calculate", "fcs_tc", "fcs_qdot", "fcs_solve_noalloc",
    "fcs_mlp_create", "fcs_mlp_destroy", "fcs_mlp_param_count", "fcs_mlp_get_params", "fcs_mlp_set_params",
    "fcs_mlp_set_norm", "fcs_mlp_forward", "fcs_mlp_loss_grad", "fcs_mlp_train",
]


class FcsError(RuntimeError):
    """A C-ABI call returned an error code (``code``, one of FCS_E_*)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


_lib = None


def load(path: str | None = None):
    """Load libfcs.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB
    if not os.path.exists(path):
        raise FcsError(f"{path} not built: run __graft_entry__.build() or python -m fcs.build")
    L = C.CDLL(path)
    L.fcs_version.restype = C.c_char_p
    L.fcs_device_count.argtypes = [C.POINTER(C.c_int)]
    L.fcs_mech_create.argtypes = [C.POINTER(MechDesc), C.POINTER(_vp)]
    L.fcs_mech_destroy.argtypes = [_vp]
    L.fcs_solver_create.argtypes = [_vp, C.POINTER(SolverOpts), C.POINTER(_vp)]
    L.fcs_solver_destroy.argtypes = [_vp]
    L.fcs_solve.argtypes = [_vp, C.c_int64, C.c_double, _dp, _dp, _dp, C.POINTER(_dp), _dp,
                            C.POINTER(_dp), _dp, _ip]
    L.fcs_solve_ex.argtypes = L.fcs_solve.argtypes + [_dp, _ip]
    L.fcs_solve_lts.argtypes = [_vp, C.c_int64, _dp, _dp, _dp, _dp, C.POINTER(_dp), _dp,
                                C.POINTER(_dp), _dp, _ip, _dp, _ip]
    L.fcs_batch_create.argtypes = [_vp, C.c_int64, C.POINTER(_vp)]
    L.fcs_batch_destroy.argtypes = [_vp]
    L.fcs_batch_upload.argtypes = [_vp, _dp, _dp, _dp, _dp, _dp]
    L.fcs_batch_run.argtypes = [_vp, C.c_double, C.POINTER(C.c_float)]
    L.fcs_batch_download.argtypes = [_vp, _dp, _dp, _dp, _ip, _dp, _ip]
    L.fcs_batch_profile.argtypes = [_vp, C.c_int]
    L.fcs_batch_stats.argtypes = [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L.fcs_batch_load_report.argtypes = [_vp, C.POINTER(C.c_int64), C.c_int32, C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int32)]
    L.fcs_eval_derivatives.argtypes = [_vp, C.c_int64, _dp, _dp, _dp]
    L.fcs_eval_jacobian.argtypes = [_vp, C.c_int64, _dp, _dp, _dp, _dp]
    L.fcs_lu_solve.argtypes = [_vp, C.c_int64, _dp, _dp]
    L.fcs_eval_attempt.argtypes = [_vp, C.c_int64, _dp, _dp, _dp, _dp, _dp]
    _fp = C.POINTER(C.c_float)
    L.fcs_batch_set_learned.argtypes = [_vp, _vp, C.c_int32]
    L.fcs_batch_learned_stats.argtypes = [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.fcs_batch_first_steps.argtypes = [_vp, _dp]
    L.fcs_batch_set_launch.argtypes = [_vp, C.c_int32, C.c_int32]
    L.fcs_batch_autotune.argtypes = [_vp, C.c_double, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    # device pointers (integers) for the arrays of the no-allocation solve
    L.fcs_solve_noalloc.argtypes = [_vp, C.c_int64, C.c_double] + [_vp] * 9 + [_dp]
    L.fcs_mlp_create.argtypes = [C.c_int32, C.POINTER(C.c_int32), C.c_uint64, C.POINTER(_vp)]
    L.fcs_mlp_destroy.argtypes = [_vp]
    L.fcs_mlp_param_count.argtypes = [_vp]
    L.fcs_mlp_get_params.argtypes = [_vp, _fp]
    L.fcs_mlp_set_params.argtypes = [_vp, _fp]
    L.fcs_mlp_set_norm.argtypes = [_vp, _fp, _fp, _fp, _fp]
    L.fcs_mlp_forward.argtypes = [_vp, C.c_int64, _fp, _fp]
    L.fcs_mlp_loss_grad.argtypes = [_vp, C.c_int64, _fp, _fp, _fp, _fp]
    L.fcs_mlp_train.argtypes = [_vp, C.c_int64, _fp, _fp, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _fp]
    for name in EXPORTS:
        getattr(L, name)
        if name != "fcs_version":
            getattr(L, name).restype = C.c_int
    L.fcs_mlp_param_count.restype = C.c_int64
    _lib = L
    return L


def check(rc, what, allow=()):
    if rc != FCS_OK and rc not in allow:
        raise FcsError(f"{what} failed with code {rc}", rc)
    return rc


def device_count() -> int:
    n = C.c_int(0)
    check(load().fcs_device_count(C.byref(n)), "fcs_device_count")
    return n.value


class Mech:
    """Device-resident mechanism tables (fcs_mech_create)."""

    def __init__(self, mech):
        self.mech = mech
        self.ns = mech.n_species
        self._desc, self._keep = make_desc(mech.arrays())
        h = _vp()
        check(load().fcs_mech_create(C.byref(self._desc), C.byref(h)), "fcs_mech_create")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.fcs_mech_destroy(self._h)
            self._h = None


class Solver:
    """fcs_solver: a mechanism plus the chemistryProperties options."""

    def __init__(self, mech: Mech, **opts):
        self.mechh = mech
        self.ns = mech.ns
        self.n = mech.ns + 1
        self.opts = make_opts(**opts)
        h = _vp()
        check(load().fcs_solver_create(mech._h, C.byref(self.opts), C.byref(h)), "fcs_solver_create")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.fcs_solver_destroy(self._h)
            self._h = None

    # ---- solve(deltaT) over host fields (an array deltaT selects the
    # local-time-step solve(scalarField deltaT), fcs_solve_lts)
    def solve(self, field, deltaT, rho=None):
        nc = field.n_cells
        T0 = np.ascontiguousarray(field.T, np.float64)
        p0 = np.ascontiguousarray(field.p, np.float64)
        rho0 = np.ascontiguousarray(field.rho(self.mechh.mech) if rho is None else rho, np.float64)
        Y0 = np.ascontiguousarray(field.Y, np.float64)
        dtc = np.array(field.deltaTChem, np.float64)
        RR = np.zeros((self.ns, nc))
        dmin = np.zeros(1)
        st = np.zeros(nc, np.int32)
        Tend = np.zeros(nc)
        na = np.zeros(nc, np.int32)
        if np.ndim(deltaT) == 0:
            rc = load().fcs_solve_ex(self._h, nc, float(deltaT), dptr(T0), dptr(p0), dptr(rho0),
                                     row_pointers(Y0), dptr(dtc), row_pointers(RR), dptr(dmin), iptr(st),
                                     dptr(Tend), iptr(na))
        else:
            dtf = np.ascontiguousarray(deltaT, np.float64)
            if dtf.shape != (nc,):
                raise ValueError("deltaT field must have one entry per cell")
            rc = load().fcs_solve_lts(self._h, nc, dptr(dtf), dptr(T0), dptr(p0), dptr(rho0),
                                      row_pointers(Y0), dptr(dtc), row_pointers(RR), dptr(dmin), iptr(st),
                                      dptr(Tend), iptr(na))
        check(rc, "fcs_solve", allow=(FCS_E_CELLS,))
        return dict(rc=rc, RR=RR, deltaTChem=dtc, deltaTMin=float(dmin[0]), status=st, Tend=Tend,
                    nAttempts=na, rho0=rho0)

    # ---- field evaluations at the current state (calculate, tc, Qdot)
    def _state(self, rho, T, p, Y):
        Y = np.ascontiguousarray(Y, np.float64).reshape(self.ns, -1)
        nc = Y.shape[1]
        arr = [np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (nc,))) for v in (rho, T, p)]
        return nc, arr, Y

    def calculate(self, rho, T, p, Y):
        """RR [ns, nCells] in kg/m^3/s at the given state (fcs_calculate)."""
        nc, (r, t, q), Y = self._state(rho, T, p, Y)
        RR = np.zeros((self.ns, nc))
        check(load().fcs_calculate(self._h, nc, dptr(r), dptr(t), dptr(q), row_pointers(Y), row_pointers(RR)),
              "fcs_calculate")
        return RR

    def tc(self, rho, T, p, Y):
        """Chemical time scale [nCells] in s (fcs_tc)."""
        nc, (r, t, q), Y = self._state(rho, T, p, Y)
        out = np.zeros(nc)
        check(load().fcs_tc(self._h, nc, dptr(r), dptr(t), dptr(q), row_pointers(Y), dptr(out)), "fcs_tc")
        return out

    def qdot(self, RR):
        """Heat release rate [nCells] in W/m^3 from RR [ns, nCells] (fcs_qdot)."""
        RR = np.ascontiguousarray(RR, np.float64).reshape(self.ns, -1)
        out = np.zeros(RR.shape[1])
        check(load().fcs_qdot(self._h, RR.shape[1], row_pointers(RR), dptr(out)), "fcs_qdot")
        return out

    # ---- kernel-level entries
    def derivatives(self, p, YT):
        YT = np.array(YT, np.float64, order="C").reshape(self.n, -1)
        nS = YT.shape[1]
        p = np.ascontiguousarray(np.broadcast_to(np.asarray(p, np.float64), (nS,)))
        out = np.zeros_like(YT)
        check(load().fcs_eval_derivatives(self._h, nS, dptr(p), dptr(YT), dptr(out)), "fcs_eval_derivatives")
        return out, YT

    def jacobian(self, p, YT):
        YT = np.array(YT, np.float64, order="C").reshape(self.n, -1)
        nS = YT.shape[1]
        p = np.ascontiguousarray(np.broadcast_to(np.asarray(p, np.float64), (nS,)))
        f = np.zeros_like(YT)
        J = np.zeros((nS, self.n, self.n))
        check(load().fcs_eval_jacobian(self._h, nS, dptr(p), dptr(YT), dptr(f), dptr(J)), "fcs_eval_jacobian")
        return f, J, YT

    def attempt(self, p, YT, dx):
        """One Rosenbrock attempt per state: returns (err, YTnew, clipped YT)."""
        YT = np.array(YT, np.float64, order="C").reshape(self.n, -1)
        nS = YT.shape[1]
        p = np.ascontiguousarray(np.broadcast_to(np.asarray(p, np.float64), (nS,)))
        dx = np.ascontiguousarray(np.broadcast_to(np.asarray(dx, np.float64), (nS,)))
        YTn = np.zeros_like(YT)
        err = np.zeros(nS)
        check(load().fcs_eval_attempt(self._h, nS, dptr(p), dptr(dx), dptr(YT), dptr(YTn), dptr(err)),
              "fcs_eval_attempt")
        return err, YTn, YT

    def lu_solve(self, A, b):
        A = np.ascontiguousarray(A, np.float64).reshape(-1, self.n, self.n)
        x = np.array(b, np.float64, order="C").reshape(-1, self.n)
        check(load().fcs_lu_solve(self._h, A.shape[0], dptr(A), dptr(x)), "fcs_lu_solve")
        return x


class Batch:
    """Device-resident batch of cells (inputs stay in HBM between runs)."""

    def __init__(self, solver: Solver, field, rho=None):
        self.s = solver
        self.nc = field.n_cells
        h = _vp()
        check(load().fcs_batch_create(solver._h, self.nc, C.byref(h)), "fcs_batch_create")
        self._h = h
        self._in = dict(
            T0=np.ascontiguousarray(field.T, np.float64), p0=np.ascontiguousarray(field.p, np.float64),
            rho0=np.ascontiguousarray(field.rho(solver.mechh.mech) if rho is None else rho, np.float64),
            Y0=np.ascontiguousarray(field.Y, np.float64), dtc=np.ascontiguousarray(field.deltaTChem, np.float64))
        i = self._in
        check(load().fcs_batch_upload(h, dptr(i["T0"]), dptr(i["p0"]), dptr(i["rho0"]), dptr(i["Y0"]),
                                      dptr(i["dtc"])), "fcs_batch_upload")

    def run(self, deltaT, timed=True) -> float:
        ms = C.c_float(0.0)
        check(load().fcs_batch_run(self._h, float(deltaT), C.byref(ms) if timed else None), "fcs_batch_run")
        return ms.value

    def profile(self, enable=True):
        """Time every step launch of later runs with a HIP event pair."""
        check(load().fcs_batch_profile(self._h, 1 if enable else 0), "fcs_batch_profile")

    def stats(self):
        """(step launches, summed step-kernel milliseconds) of the last run."""
        n, ms = C.c_int64(0), C.c_double(0.0)
        check(load().fcs_batch_stats(self._h, C.byref(n), C.byref(ms)), "fcs_batch_stats")
        return n.value, ms.value

    def load_report(self):
        """(list length read by each step launch of the last profiled eager
        run, step-launch grid)."""
        n, g = C.c_int32(0), C.c_int32(0)
        check(load().fcs_batch_load_report(self._h, None, 0, C.byref(n), C.byref(g)), "fcs_batch_load_report")
        out = np.zeros(n.value, np.int64)
        check(load().fcs_batch_load_report(self._h, out.ctypes.data_as(C.POINTER(C.c_int64)), n.value, C.byref(n),
                                           C.byref(g)), "fcs_batch_load_report")
        return out, g.value

    def set_launch(self, grid=0, graph_launches=0):
        """Step-kernel grid (0: occupancy x CUs) and step launches per
        captured hipGraph (0: eager launches)."""
        check(load().fcs_batch_set_launch(self._h, int(grid), int(graph_launches)), "fcs_batch_set_launch")

    def autotune(self, deltaT):
        """Times the candidate launch shapes on the uploaded cells and keeps
        the fastest; returns (grid, graph_launches)."""
        g, n = C.c_int32(0), C.c_int32(0)
        check(load().fcs_batch_autotune(self._h, float(deltaT), C.byref(g), C.byref(n)), "fcs_batch_autotune")
        return g.value, n.value

    def set_learned(self, mlp, mode="verify"):
        """Learned seed on later runs: mode "off", "verify" (fallback cells
        integrate exactly as without the network) or "seed" (fallback cells
        start from the network's first sub-step)."""
        m = {"off": 0, "verify": 1, "seed": 2}[mode]
        check(load().fcs_batch_set_learned(self._h, None if mlp is None else mlp._h, m), "fcs_batch_set_learned")
        self._mlp = mlp     # the batch keeps the network's pointer

    def learned_stats(self):
        """(cells given a proposal, proposals verified, proposals accepted)
        of the last run."""
        a, b, c = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        check(load().fcs_batch_learned_stats(self._h, C.byref(a), C.byref(b), C.byref(c)),
              "fcs_batch_learned_stats")
        return a.value, b.value, c.value

    def first_steps(self):
        out = np.zeros(self.nc)
        check(load().fcs_batch_first_steps(self._h, dptr(out)), "fcs_batch_first_steps")
        return out

    def download(self):
        ns, nc = self.s.ns, self.nc
        RR = np.zeros((ns, nc))
        dtc = np.zeros(nc)
        dmin = np.zeros(1)
        st = np.zeros(nc, np.int32)
        Tend = np.zeros(nc)
        na = np.zeros(nc, np.int32)
        check(load().fcs_batch_download(self._h, dptr(RR), dptr(dtc), dptr(dmin), iptr(st), dptr(Tend),
                                        iptr(na)), "fcs_batch_download")
        return dict(RR=RR, deltaTChem=dtc, deltaTMin=float(dmin[0]), status=st, Tend=Tend, nAttempts=na)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.fcs_batch_destroy(self._h)
            self._h = None


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """The learned seed's network on the GPU (fcs_mlp_*): BF16 MFMA forward
    and backward kernels, FP32 master weights, Adam."""

    def __init__(self, widths, seed=0):
        self.widths = [int(w) for w in widths]
        wa = (C.c_int32 * len(self.widths))(*self.widths)
        h = _vp()
        check(load().fcs_mlp_create(len(self.widths) - 1, wa, int(seed), C.byref(h)), "fcs_mlp_create")
        self._h = h
        self.n_params = int(load().fcs_mlp_param_count(h))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.fcs_mlp_destroy(self._h)
            self._h = None

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        check(load().fcs_mlp_get_params(self._h, _fptr(p)), "fcs_mlp_get_params")
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.shape == (self.n_params,)
        check(load().fcs_mlp_set_params(self._h, _fptr(p)), "fcs_mlp_set_params")

    def set_norm(self, in_mean, in_std, out_mean, out_std):
        a, b, c, d = (_f32(x) for x in (in_mean, in_std, out_mean, out_std))
        self.norm = (a, b, c, d)
        check(load().fcs_mlp_set_norm(self._h, _fptr(a), _fptr(b), _fptr(c), _fptr(d)), "fcs_mlp_set_norm")

    def forward(self, X):
        X = _f32(X)
        Y = np.zeros((X.shape[0], self.widths[-1]), np.float32)
        check(load().fcs_mlp_forward(self._h, X.shape[0], _fptr(X), _fptr(Y)), "fcs_mlp_forward")
        return Y

    def loss_grad(self, X, T):
        X, T = _f32(X), _f32(T)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.n_params, np.float32)
        check(load().fcs_mlp_loss_grad(self._h, X.shape[0], _fptr(X), _fptr(T), _fptr(loss), _fptr(g)),
              "fcs_mlp_loss_grad")
        return float(loss[0]), g

    def train(self, X, T, epochs, batch=256, lr=1e-3, seed=0):
        X, T = _f32(X), _f32(T)
        hist = np.zeros(epochs, np.float32)
        check(load().fcs_mlp_train(self._h, X.shape[0], _fptr(X), _fptr(T), int(epochs), int(batch), float(lr),
                                   int(seed), _fptr(hist)), "fcs_mlp_train")
        return hist


def wavefront_loads(list_len, grid):
    """Attempts each step-kernel wavefront made over a run, from the list
    length n of every step launch (Batch.load_report): the launch strides
    its n cells over the grid, so wavefront w takes cells w, w + grid, ...,
    ceil((n - w) / grid) of them (0 when n <= w)."""
    n = np.asarray(list_len, np.int64)[:, None]
    w = np.arange(int(grid), dtype=np.int64)[None, :]
    return np.maximum(0, (n - w + grid - 1) // grid).sum(axis=0)


def load_summary(list_len, grid):
    """Per-wavefront load figures of a run: attempts per wavefront (min,
    mean, max), the launches whose list holds fewer cells than the grid has
    wavefronts (the tail, where wavefronts idle) and their share of the
    launches and of the attempts."""
    n = np.asarray(list_len, np.int64)
    loads = wavefront_loads(n, grid)
    tail = n < grid
    return {"grid": int(grid), "launches": int(n.size), "attempts": int(n.sum()),
            "attempts_per_wavefront": {"min": int(loads.min()), "mean": float(loads.mean()),
                                       "max": int(loads.max())},
            "tail_launches": int(tail.sum()), "tail_attempt_share": float(n[tail].sum() / max(n.sum(), 1)),
            "mean_busy_wavefronts_per_launch": float(np.minimum(n, grid).mean()) if n.size else 0.0}


def equilibrator(solver: "Solver", deltaT=4e-3):
    """An ``equilibrate`` callable for the counterflow field generator: it
    integrates a Field for deltaT (default 4 ms, long past the initial
    transient of the synthetic 53-species mechanism) on the GPU and returns
    the end states (T from Tend, Y = Y0 + RR deltaT / rho0, clipped and
    renormalised)."""
    from .mechanism import Field

    def run(field):
        r = solver.solve(field, deltaT)
        if r["rc"] != FCS_OK:
            raise FcsError(f"equilibration failed with code {r['rc']}", r["rc"])
        Y = np.maximum(field.Y + r["RR"] * deltaT / r["rho0"][None, :], 0.0)
        return Field(r["Tend"].copy(), field.p.copy(), Y / Y.sum(axis=0), field.deltaTChem.copy())
    return run
