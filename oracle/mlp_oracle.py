"""CPU oracle of the learned seed's network (test infrastructure only).

Restates, in NumPy, what the gfx950 kernels of fcs_mlp compute: inputs
normalised with (x - mean) / std, per layer Z = H W + b with SiLU on hidden
layers, outputs de-normalised; loss = mean squared error of the normalised
outputs; gradients by backpropagation.  Parameters use the flat padded
layout of fcs_mlp_get_params (per layer W [pad16(in)][pad16(out)] then
b [pad16(out)]).

``bf16=True`` rounds every matrix-product operand to BF16 (round to nearest
even), as the MFMA operand loads do, and accumulates in float64; with
``bf16=False`` the whole computation is float64.  The float64 gradient is
pinned by central finite differences in tests/test_mlp_oracle.py.
"""
from __future__ import annotations

import numpy as np


def pad16(x):
    return (x + 15) // 16 * 16


def to_bf16(x):
    """float32 -> nearest-even bfloat16, returned as float32 values."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = ((u & 0x7F800000) == 0x7F800000) & ((u & 0x007FFFFF) != 0)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    r = np.where(nan, (u | 0x00400000) & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


class MlpOracle:
    def __init__(self, widths):
        self.w = list(widths)
        self.wp = [pad16(v) for v in widths]
        self.offW, self.offB = [], []
        off = 0
        for l in range(len(widths) - 1):
            self.offW.append(off)
            off += self.wp[l] * self.wp[l + 1]
            self.offB.append(off)
            off += self.wp[l + 1]
        self.n_params = off
        self.in_mean = np.zeros(widths[0])
        self.in_std = np.ones(widths[0])
        self.out_mean = np.zeros(widths[-1])
        self.out_std = np.ones(widths[-1])

    def set_norm(self, a, b, c, d):
        self.in_mean, self.in_std, self.out_mean, self.out_std = (np.asarray(v, np.float64) for v in (a, b, c, d))

    def unpack(self, p):
        p = np.asarray(p, np.float64)
        Ws, bs = [], []
        for l in range(len(self.w) - 1):
            W = p[self.offW[l]:self.offW[l] + self.wp[l] * self.wp[l + 1]].reshape(self.wp[l], self.wp[l + 1])
            Ws.append(W[:self.w[l], :self.w[l + 1]])
            bs.append(p[self.offB[l]:self.offB[l] + self.w[l + 1]])
        return Ws, bs

    def pack(self, Ws, bs):
        p = np.zeros(self.n_params)
        for l, (W, b) in enumerate(zip(Ws, bs)):
            blk = np.zeros((self.wp[l], self.wp[l + 1]))
            blk[:self.w[l], :self.w[l + 1]] = W
            p[self.offW[l]:self.offW[l] + blk.size] = blk.ravel()
            p[self.offB[l]:self.offB[l] + self.w[l + 1]] = b
        return p

    @staticmethod
    def _mm(A, B, bf16):
        if bf16:
            A, B = to_bf16(A).astype(np.float64), to_bf16(B).astype(np.float64)
        return A @ B

    def _forward(self, p, X, bf16):
        Ws, bs = self.unpack(p)
        H = (np.asarray(X, np.float64) - self.in_mean) / self.in_std
        if bf16:
            H = H.astype(np.float32).astype(np.float64)
        Hs, Zs = [H], []
        for l, (W, b) in enumerate(zip(Ws, bs)):
            Z = self._mm(Hs[-1], W, bf16) + b
            Zs.append(Z)
            Hs.append(Z / (1.0 + np.exp(-Z)) if l + 1 < len(Ws) else Z)
        return Hs, Zs

    def forward(self, p, X, bf16=True):
        Hs, _ = self._forward(p, X, bf16)
        return Hs[-1] * self.out_std + self.out_mean

    def loss_grad(self, p, X, T, bf16=True):
        """(loss, flat gradient) of the mean squared error of normalised outputs."""
        Ws, _ = self.unpack(p)
        Hs, Zs = self._forward(p, X, bf16)
        Tn = (np.asarray(T, np.float64) - self.out_mean) / self.out_std
        n, m = Tn.shape
        d = Hs[-1] - Tn
        loss = float((d * d).sum() / (n * m))
        dcur = 2.0 * d / (n * m)
        gW, gb = [None] * len(Ws), [None] * len(Ws)
        for l in range(len(Ws) - 1, -1, -1):
            if l + 1 < len(Ws):
                s = 1.0 / (1.0 + np.exp(-Zs[l]))
                dcur = dcur * s * (1.0 + Zs[l] * (1.0 - s))
            gW[l] = self._mm(Hs[l].T, dcur, bf16)
            gb[l] = dcur.sum(axis=0)
            if l > 0:
                dcur = self._mm(dcur, Ws[l].T, bf16)
        return loss, self.pack(gW, gb)
