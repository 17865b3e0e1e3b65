"""Learned seed of the integrator on the GPU (configuration 3 on small
counterflow fields): fallback cells are bit-identical to the plain
integrator, accepted proposals meet the parity bar against it, and a
trained network saves attempts on the near-equilibrium product cells."""
import numpy as np
import pytest

import fcs.mechanism as M

pytestmark = pytest.mark.gpu

DT = 1e-6


@pytest.fixture(scope="module")
def fcslib():
    from fcs import lib
    lib.load()
    assert lib.device_count() >= 1
    return lib


@pytest.fixture(scope="module")
def setup(fcslib, mech1):
    s = fcslib.Solver(fcslib.Mech(mech1), Treact=500.0)
    eq = fcslib.equilibrator(s)
    fields = {seed: M.counterflow_field(mech1, 48, seed=seed, equilibrate=eq) for seed in (11, 12, 13)}
    return s, fields


def end_Y(field, r, rho0):
    return field.Y + r["RR"] * DT / rho0[None, :]


def close(field, a, b, rho0):
    Ya, Yb = end_Y(field, a, rho0), end_Y(field, b, rho0)
    d = np.abs(Ya - Yb)
    okY = ((d <= 1e-6) | (d <= 1e-4 * np.abs(Yb))).all(axis=0)
    return okY & (np.abs(a["Tend"] - b["Tend"]) <= 1e-3)


def run(fcslib, s, field, mlp=None, mode="verify"):
    b = fcslib.Batch(s, field)
    if mlp is not None:
        b.set_learned(mlp, mode)
    b.run(DT)
    r = b.download()
    return r, b.learned_stats()


def test_untrained_network_leaves_fallback_cells_bit_identical(fcslib, setup, mech1):
    s, fields = setup
    f = fields[13]
    net = fcslib.Mlp([56, 128, 128, 128, 55], seed=9)
    plain, _ = run(fcslib, s, f)
    lrn, (npred, nver, nacc) = run(fcslib, s, f, net)
    react = f.T > 500.0
    assert npred == react.sum() and nacc <= nver <= npred
    acc = (lrn["status"] & 8) != 0
    assert acc.sum() == nacc and (lrn["nAttempts"][acc] == 1).all()
    fb = react & ~acc
    for k in ("RR", "Tend", "deltaTChem", "nAttempts", "status"):   # bit-identical fallback
        a, b = lrn[k], plain[k]
        if a.ndim == 2:
            a, b = a[:, fb], b[:, fb]
        else:
            a, b = a[fb], b[fb]
        np.testing.assert_array_equal(a, b)
    rho0 = f.rho(mech1)
    assert close(f, lrn, plain, rho0)[react].all()
    assert (lrn["status"][~react] == 1).all()


def test_trained_network_saves_attempts_on_product_cells(fcslib, setup, mech1):
    from fcs import learned
    s, fields = setup
    Xs, Ts = zip(*(learned.samples(s, fields[k], DT) for k in (11, 12)))
    X, T = np.concatenate(Xs), np.concatenate(Ts)
    net, hist = learned.train(X, T, epochs=40, batch=256, lr=2e-3, seed=1)
    assert hist[-1] < 0.5 * hist[0]
    f = fields[13]
    plain, _ = run(fcslib, s, f)
    lrn, (npred, nver, nacc) = run(fcslib, s, f, net)
    prod = f.cls == 1
    acc = (lrn["status"] & 8) != 0
    assert acc[prod].sum() >= 0.5 * prod.sum(), (acc[prod].sum(), prod.sum(), nver, nacc)
    # attempts on product cells, counting one verification step for every
    # verified cell that fell back: clearly fewer than the plain integrator's
    att_plain = plain["nAttempts"][prod].sum()
    att_lrn = lrn["nAttempts"][prod].sum() + (nver - nacc)
    assert att_lrn <= 0.75 * att_plain, (att_lrn, att_plain)
    # the gate keeps the flame-front cells away from the verification step
    flame = f.cls == 2
    assert nver <= prod.sum() + 0.1 * flame.sum(), (nver, prod.sum(), flame.sum())
    rho0 = f.rho(mech1)
    react = f.T > 500.0
    assert close(f, lrn, plain, rho0)[react].all()
    # seed mode: fallback cells start from the proposed first sub-step, a
    # different step sequence; its error against a tight-tolerance solve
    # (relTol 1e-10) stays at the plain integrator's level
    sd, _ = run(fcslib, s, f, net, mode="seed")
    assert ((sd["status"] & 7) == (plain["status"] & 7)).all()
    idx = np.concatenate([np.nonzero(flame)[0][:12], np.nonzero(prod)[0][:12]])
    sub = f.subset(idx)
    tight = fcslib.Solver(fcslib.Mech(mech1), Treact=500.0, relTol=1e-10, absTol=1e-14).solve(sub, DT)
    Yt = end_Y(sub, tight, tight["rho0"])

    def err(r):
        Y = end_Y(sub, {"RR": r["RR"][:, idx]}, tight["rho0"])
        eY = (np.abs(Y - Yt) / (1e-8 + 1e-4 * np.abs(Yt))).max()
        eT = (np.abs(r["Tend"][idx] - tight["Tend"]) / (1e-4 * tight["Tend"])).max()
        return max(eY, eT)
    e_plain, e_seed, e_lrn = err(plain), err(sd), err(lrn)
    assert e_seed <= max(3.0 * e_plain, 1.0) and e_lrn <= max(3.0 * e_plain, 1.0), (e_plain, e_seed, e_lrn)
