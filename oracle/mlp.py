"""CPU oracle of the surrogate network (numpy).

TEST INFRASTRUCTURE ONLY: the checker of the GPU network's forward pass, gradients and Adam step.

It restates the arithmetic declared in include/spp_surrogate.h: weights and activations rounded to
bfloat16 (round to nearest even) where they enter a matrix product, FP32 accumulation, FP32 biases,
pre-activations, loss, gradients and Adam moments.  `exact=True` runs the same network in float64
without any rounding, which the finite-difference check pins.
"""
from __future__ import annotations

import numpy as np


def bf16_round(a):
    """float32 -> nearest-even bfloat16, returned as float32 (NaN kept quiet)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def layer_dims(n_in, hidden, n_hidden, n_out):
    return [n_in] + [hidden] * n_hidden + [n_out]


def unpack(params, dims):
    """Flat vector -> [(W [k][j], b [j])] in the order of include/spp_surrogate.h."""
    out, o = [], 0
    for din, dout in zip(dims[:-1], dims[1:]):
        W = params[o:o + din * dout].reshape(din, dout)
        o += din * dout
        b = params[o:o + dout]
        o += dout
        out.append((W, b))
    return out


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


class MLPOracle:
    def __init__(self, dims, params, exact=False):
        self.dims = list(dims)
        self.exact = exact
        self.dt = np.float64 if exact else np.float32
        self.params = np.asarray(params, dtype=self.dt).copy()

    def _r(self, a):
        return a.astype(self.dt) if self.exact else bf16_round(a)

    def forward(self, x, keep=False):
        layers = unpack(self.params, self.dims)
        H = self._r(np.asarray(x, dtype=self.dt))
        Hs, Zs = [H], []
        for l, (W, b) in enumerate(layers):
            Z = (H @ self._r(W)).astype(self.dt) + b
            if l < len(layers) - 1:
                Zs.append(Z)
                H = self._r(Z * _sig(Z))
                Hs.append(H)
            else:
                Y = Z
        if keep:
            return Y, Hs, Zs
        return Y

    def loss_grad(self, x, t, w=None):
        layers = unpack(self.params, self.dims)
        Y, Hs, Zs = self.forward(x, keep=True)
        B, n_out = Y.shape
        wv = np.ones(n_out, dtype=self.dt) if w is None else np.asarray(w, dtype=self.dt)
        d = Y - np.asarray(t, dtype=self.dt)
        loss = float((wv * d * d).sum() / (B * n_out))
        dZ = (2.0 * wv * d / (B * n_out)).astype(self.dt)
        grads = [None] * len(layers)
        for l in range(len(layers) - 1, -1, -1):
            W, _ = layers[l]
            dZr = self._r(dZ)
            gW = (Hs[l].T @ dZr).astype(self.dt)
            gb = dZ.sum(axis=0).astype(self.dt)
            grads[l] = (gW, gb)
            if l == 0:
                break
            dH = (dZr @ self._r(W).T).astype(self.dt)
            z = Zs[l - 1]
            s = _sig(z)
            dZ = (dH * (s * (1 + z * (1 - s)))).astype(self.dt)
        flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
        return loss, flat

    def adam_step(self, grad, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        """One Adam update in FP32 (t counts from 1); returns the new (m, v)."""
        f = np.float32
        g = grad.astype(f)
        m = f(b1) * m + f(1 - b1) * g
        v = f(b2) * v + f(1 - b2) * g * g
        c1 = f(1) - f(b1) ** f(t)
        c2 = f(1) - f(b2) ** f(t)
        self.params = (self.params.astype(f) - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))).astype(f)
        return m, v
