"""FP64 numpy reference of the warm-start network -- TEST INFRASTRUCTURE ONLY.

The same three-layer ReLU perceptron as superdca_amd/csrc/sdca_mlp.hip with the same
parameter layout (W1 [H][Din], b1 [H], W2 [H][H], b2 [H], W3 [Dout][H], b3 [Dout]) and
loss (1/sum(mask)) sum mask (y - t)^2.  Its gradient is pinned by central finite
differences in tests/test_oracle.py; the GPU kernels are checked against it.
"""
from __future__ import annotations

import numpy as np


def unpack(p: np.ndarray, din: int, dh: int, dout: int):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + dh * din].reshape(dh, din); o += dh * din
    b1 = p[o:o + dh]; o += dh
    W2 = p[o:o + dh * dh].reshape(dh, dh); o += dh * dh
    b2 = p[o:o + dh]; o += dh
    W3 = p[o:o + dout * dh].reshape(dout, dh); o += dout * dh
    b3 = p[o:o + dout]; o += dout
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def forward(p, x, din, dh, dout):
    W1, b1, W2, b2, W3, b3 = unpack(p, din, dh, dout)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3.T + b3, (h1, h2)


def loss_grad(p, x, t, mask, din, dh, dout):
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    m = np.ones_like(t) if mask is None else np.asarray(mask, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(p, din, dh, dout)
    y, (h1, h2) = forward(p, x, din, dh, dout)
    cnt = m.sum()
    d = (y - t) * m
    loss = (d * d).sum() / cnt
    dy = 2.0 * d / cnt
    gW3 = dy.T @ h2
    gb3 = dy.sum(axis=0)
    dz2 = (dy @ W3) * (h2 > 0)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2 @ W2) * (h1 > 0)
    gW1 = dz1.T @ x
    gb1 = dz1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
