"""CPU oracle of the radiation-surrogate MLP (test infrastructure only).

The network of l-game_amd/csrc/lgame_mlp.hip in numpy: n_in -> hidden x n_hidden (exact GELU) ->
n_out, parameters in the order W_1 [out][in], b_1, W_2, b_2, ..., inputs and targets normalised
with (value - mean) / scale, loss = mean squared error of the normalised outputs.

With ``bf16=True`` the values are rounded to bfloat16 where the device stores them in bfloat16
(normalised inputs, weights, pre-activations, activations, back-propagated gradients) and all
arithmetic is FP64; with ``bf16=False`` everything is FP64 (the form the finite-difference check
pins).  There is no reference counterpart: the reference calls RTE+RRTMGP instead.
"""
from __future__ import annotations

import math

import numpy as np

try:
    from scipy.special import erf as _erf
except ImportError:  # pragma: no cover
    _erf = np.vectorize(math.erf)


def bf16(a):
    """Round float32 values to bfloat16 (nearest even) and return them as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


class MlpOracle:
    def __init__(self, n_in, hidden, n_hidden, n_out, params, x_mean=None, x_scale=None, y_mean=None, y_scale=None,
                 bf16=True):
        self.sizes = [n_in] + [hidden] * n_hidden + [n_out]
        self.bf = bf16
        self.W, self.b = [], []
        p = np.asarray(params, dtype=np.float64)
        off = 0
        for i, o in zip(self.sizes[:-1], self.sizes[1:]):
            self.W.append(p[off:off + i * o].reshape(o, i))
            off += i * o
            self.b.append(p[off:off + o])
            off += o
        assert off == p.size
        self.xm = np.zeros(n_in) if x_mean is None else np.asarray(x_mean, np.float64)
        self.xs = np.ones(n_in) if x_scale is None else np.asarray(x_scale, np.float64)
        self.ym = np.zeros(n_out) if y_mean is None else np.asarray(y_mean, np.float64)
        self.ys = np.ones(n_out) if y_scale is None else np.asarray(y_scale, np.float64)

    def _r(self, a):
        return bf16(a) if self.bf else a

    def _forward(self, x):
        xn = self._r((np.asarray(x, np.float32) - self.xm.astype(np.float32)) / self.xs.astype(np.float32))
        acts, zs = [xn], []
        a = xn
        L = len(self.W)
        for l in range(L):
            z = a @ self._r(self.W[l]).T + self.b[l]
            if l < L - 1:
                zs.append(self._r(z))
                a = self._r(gelu(z))
                acts.append(a)
            else:
                a = z
        return a, acts, zs

    def forward(self, x):
        yn, _, _ = self._forward(x)
        return yn * self.ys + self.ym

    def loss_and_grad(self, x, y):
        yn, acts, zs = self._forward(x)
        t = (np.asarray(y, np.float32) - self.ym.astype(np.float32)) / self.ys.astype(np.float32)
        n, no = yn.shape
        e = yn - t
        loss = float(np.mean(e * e))
        d = self._r(2.0 * e / (n * no))
        gW, gb = [None] * len(self.W), [None] * len(self.W)
        for l in range(len(self.W) - 1, -1, -1):
            gW[l] = d.T @ acts[l]
            gb[l] = d.sum(axis=0)
            if l > 0:
                d = self._r((d @ self._r(self.W[l])) * gelu_grad(zs[l - 1]))
        g = np.concatenate([np.concatenate([gw.ravel(), gbb]) for gw, gbb in zip(gW, gb)])
        return loss, g
