"""Pins the CPU oracle of the learned seed's network: BF16 rounding on
known values, and the float64 backpropagation gradient against central
finite differences of the loss."""
import numpy as np

from mlp_oracle import MlpOracle, to_bf16


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 1.0 + 2.0 ** -9, -2.5, 3.0e38, np.inf], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -7, 1.0, -2.5, 3.0040553e38, np.inf], np.float32)
    np.testing.assert_array_equal(to_bf16(x), want)
    assert np.isnan(to_bf16(np.array([np.nan], np.float32)))[0]


def test_pack_unpack_round_trip():
    o = MlpOracle([5, 7, 3])
    rng = np.random.default_rng(0)
    Ws = [rng.standard_normal((5, 7)), rng.standard_normal((7, 3))]
    bs = [rng.standard_normal(7), rng.standard_normal(3)]
    p = o.pack(Ws, bs)
    assert p.size == 16 * 16 + 16 + 16 * 16 + 16
    W2, b2 = o.unpack(p)
    for a, b in zip(Ws + bs, W2 + b2):
        np.testing.assert_array_equal(a, b)


def test_gradient_matches_finite_differences():
    o = MlpOracle([6, 9, 9, 4])
    rng = np.random.default_rng(1)
    o.set_norm(rng.standard_normal(6), rng.uniform(0.5, 2, 6), rng.standard_normal(4), rng.uniform(0.5, 2, 4))
    Ws = [0.5 * rng.standard_normal((6, 9)), 0.5 * rng.standard_normal((9, 9)), 0.5 * rng.standard_normal((9, 4))]
    bs = [0.1 * rng.standard_normal(9), 0.1 * rng.standard_normal(9), 0.1 * rng.standard_normal(4)]
    p = o.pack(Ws, bs)
    X = rng.standard_normal((11, 6))
    T = rng.standard_normal((11, 4))
    loss, g = o.loss_grad(p, X, T, bf16=False)
    assert loss > 0
    # every real (unpadded) parameter, central differences
    Wm, bm = o.unpack(np.arange(o.n_params, dtype=np.float64))
    idx = np.concatenate([w.ravel() for w in Wm] + [b for b in bm]).astype(int)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (o.loss_grad(pp, X, T, bf16=False)[0] - o.loss_grad(pm, X, T, bf16=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # padding entries carry no gradient
    mask = np.ones(o.n_params, bool)
    mask[idx] = False
    assert (g[mask] == 0).all()
