"""The learned seed's network on the GPU (BF16 MFMA forward/backward,
FP32 accumulation) against the CPU oracle (oracle/mlp_oracle.py) at the
configuration-3 shape 56 -> 128 -> 128 -> 128 -> 55."""
import numpy as np
import pytest

from mlp_oracle import MlpOracle

pytestmark = pytest.mark.gpu

WIDTHS = [56, 128, 128, 128, 55]


@pytest.fixture(scope="module")
def fcslib():
    from fcs import lib
    lib.load()
    assert lib.device_count() >= 1
    return lib


def make(fcslib, seed=3):
    rng = np.random.default_rng(seed)
    net = fcslib.Mlp(WIDTHS, seed=seed)
    o = MlpOracle(WIDTHS)
    norm = (rng.standard_normal(56), rng.uniform(0.5, 2.0, 56), rng.standard_normal(55), rng.uniform(0.5, 2.0, 55))
    net.set_norm(*norm)
    o.set_norm(*(np.asarray(v, np.float32).astype(np.float64) for v in norm))
    return net, o, rng


def test_forward_matches_oracle(fcslib):
    net, o, rng = make(fcslib)
    p = net.params().astype(np.float64)
    assert p.size == o.n_params
    X = rng.standard_normal((200, 56)) * 2.0
    Yg = net.forward(X)
    Yo = o.forward(p, X.astype(np.float32), bf16=True)
    scale = np.abs(Yo).max(axis=0) + 1e-3
    assert (np.abs(Yg - Yo) / scale).max() <= 2e-3
    # and far from the unrounded float64 network only by BF16 rounding
    Y64 = o.forward(p, X.astype(np.float32), bf16=False)
    assert (np.abs(Yg - Y64) / scale).max() <= 5e-2


def test_gradient_matches_oracle(fcslib):
    net, o, rng = make(fcslib, seed=4)
    p = net.params().astype(np.float64)
    X = rng.standard_normal((96, 56))
    T = rng.standard_normal((96, 55))
    lg, gg = net.loss_grad(X, T)
    lo, go = o.loss_grad(p, X.astype(np.float32), T.astype(np.float32), bf16=True)
    assert lg == pytest.approx(lo, rel=2e-3)
    # per layer, relative to the layer's gradient norm
    for l in range(len(WIDTHS) - 1):
        a, b = o.offW[l], o.offB[l] + o.wp[l + 1]
        rel = np.linalg.norm(gg[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
        assert rel <= 1e-2, (l, rel)
    assert (gg[go == 0] == 0).all()    # padding parameters get no gradient


def test_training_lowers_the_loss(fcslib):
    net, o, rng = make(fcslib, seed=5)
    net.set_norm(np.zeros(56), np.ones(56), np.zeros(55), np.ones(55))
    X = rng.standard_normal((4096, 56))
    A = rng.standard_normal((56, 55)) / 8.0
    T = np.tanh(X @ A)
    l0, _ = net.loss_grad(X, T)
    hist = net.train(X, T, epochs=30, batch=256, lr=2e-3, seed=1)
    l1, _ = net.loss_grad(X, T)
    assert hist[-1] < hist[0] and l1 < 0.1 * l0, (l0, l1, hist[[0, -1]])
