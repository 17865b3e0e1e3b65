"""Training data and training of the learned seed (configuration 3).

Samples come from the plain GPU integrator, started with deltaTChem equal
to the flow step: for every reacting cell of a field, the inputs
[T, ln p, Y_0..Y_ns-1, ln deltaT] and the targets [ln h, z_0..z_ns-1, z_T]
where h is the first accepted sub-step (deltaT itself when the cell can take
the whole step at once) and z_k = asinh(d_k / (absTol + relTol |y0_k|)) the
tolerance-scaled, compressed end-state increments.  The network
(fcs.lib.Mlp, BF16 MFMA kernels) is trained on them with Adam; the
integrator then uses it through Batch.set_learned.
"""
from __future__ import annotations

import numpy as np

from . import lib
from .mechanism import Field


def samples(solver: "lib.Solver", field: Field, deltaT: float, absTol=1e-8, relTol=1e-4):
    """(X, T) float32 arrays of the cells of `field` that the plain
    integrator finishes with at least one accepted sub-step."""
    field = Field(field.T, field.p, field.Y, np.full(field.n_cells, deltaT), field.cls)
    b = lib.Batch(solver, field)
    b.run(deltaT)
    r = b.download()
    first = b.first_steps()
    ok = (r["status"] == 0) & (first > 0)
    rho0 = field.rho(solver.mechh.mech)
    dY = r["RR"] * deltaT / rho0[None, :]
    X = np.concatenate([field.T[None, :], np.log(field.p)[None, :], np.maximum(field.Y, 0.0),
                        np.full((1, field.n_cells), np.log(deltaT))], axis=0).T
    Y0 = np.maximum(field.Y, 0.0)
    zY = np.arcsinh(dY / (absTol + relTol * Y0))
    zT = np.arcsinh((r["Tend"] - field.T) / (absTol + relTol * field.T))
    T = np.concatenate([np.log(np.where(first > 0, first, 1.0))[None, :], zY, zT[None, :]], axis=0).T
    return X[ok].astype(np.float32), T[ok].astype(np.float32)


def normaliser(X, T, lnh_scale=None):
    """Per-column mean / std (std floored so constant columns stay finite).
    The first target, ln h, is bimodal on flame fields (ln(h/deltaT) = 0 for
    cells that take the flow step at once, about -30 for flame-front cells),
    so its spread (~13) makes an error of a few tenths of an e-fold cheap in
    the loss while the verification gate (h >= deltaT/2) hinges on exactly
    that; lnh_scale, when given, caps its scale at that many e-folds (with a
    few thousand samples that keeps the gate open for product cells; the
    2 M-sample training of the benchmark does not need it)."""
    def ms(A):
        m = A.mean(axis=0, dtype=np.float64)
        s = A.std(axis=0, dtype=np.float64)
        s = np.where(s > 1e-12 * (np.abs(m) + 1e-30), s, np.maximum(np.abs(m), 1.0))
        return m, np.maximum(s, 1e-30)
    xm, xs = ms(X)
    tm, ts = ms(T)
    if lnh_scale:
        ts[0] = min(ts[0], lnh_scale)
    return xm, xs, tm, ts


def train(X, T, hidden=(128, 128, 128), epochs=40, batch=512, lr=2e-3, seed=0):
    """A fresh network of widths [in, *hidden, out] trained on (X, T)."""
    net = lib.Mlp([X.shape[1], *hidden, T.shape[1]], seed=seed)
    net.set_norm(*normaliser(X, T))
    hist = net.train(X, T, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return net, hist
