"""FP64 numpy reference of the warm-start network and its pair features -- TEST INFRASTRUCTURE ONLY.

The same three-layer ReLU perceptron as superdca_amd/csrc/sdca_mlp.hip with the same
parameter layout (W1 [H][Din], b1 [H], W2 [H][H], b2 [H], W3 [Dout][H], b3 [Dout]) and
loss (1/sum(mask)) sum mask (y - t)^2, in two forms: exact FP64 (loss_grad, pinned by
central finite differences in tests/test_oracle.py) and with the device's BF16 rounding
points (loss_grad_bf16; its rounding function is pinned against torch's bfloat16
conversion, its gradient against the FP64 one).  The pair features' mean-field block
(mean_field_block, a full 2(q-1) x 2(q-1) inverse) is pinned by the q = 2 closed form,
independence and the transpose symmetry of swapping the pair.  The GPU kernels are
checked against these.
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


# ---- the BF16 network and the pair features

def bf16(x) -> np.ndarray:
    """FP32 -> BF16 -> FP32 with round to nearest even (NaN stays NaN), as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(a.shape)


def forward_bf16(p, x, din, dh, dout):
    """The device network's forward: BF16 input, weights and hidden activations, products
    and biases in FP64 (the device accumulates in FP32)."""
    W1, b1, W2, b2, W3, b3 = unpack(p, din, dh, dout)
    xb = bf16(x).astype(np.float64)
    w1, w2, w3 = (bf16(W).astype(np.float64) for W in (W1, W2, W3))
    h1 = bf16(np.maximum(xb @ w1.T + b1, 0.0)).astype(np.float64)
    h2 = bf16(np.maximum(h1 @ w2.T + b2, 0.0)).astype(np.float64)
    return h2 @ w3.T + b3, (xb, h1, h2, w1, w2, w3)


def loss_grad_bf16(p, x, t, mask, din, dh, dout):
    """The device network's loss and gradient: output gradients rounded to BF16 where the
    device rounds them (dY before the products, dZ2 before the next layer's products, dZ1
    before the first layer's weight product); bias gradients from the unrounded values."""
    t = np.asarray(t, dtype=np.float64)
    m = np.ones_like(t) if mask is None else np.asarray(mask, dtype=np.float64)
    y, (xb, h1, h2, w1, w2, w3) = forward_bf16(p, x, din, dh, dout)
    cnt = m.sum()
    d = (y - t) * m
    loss = (d * d).sum() / cnt
    dy = 2.0 * d / cnt
    dyb = bf16(dy).astype(np.float64)
    gW3 = dyb.T @ h2
    gb3 = dy.sum(axis=0)
    dz2 = (dyb @ w3) * (h2 > 0)
    dz2b = bf16(dz2).astype(np.float64)
    gW2 = dz2b.T @ h1
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2b @ w2) * (h1 > 0)
    dz1b = bf16(dz1).astype(np.float64)
    gW1 = dz1b.T @ xb
    gb1 = dz1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def mean_field_block(N, B: int, q: int, pseudocount: float = 0.5) -> np.ndarray:
    """Mean-field coupling estimate of one locus pair from its q x q counts N[s, a] (s at r,
    a at j): J = -(C^-1)_rj of the pair's 2(q-1) x 2(q-1) connected-correlation matrix over
    the first q-1 states of each locus (the last is the reference state), frequencies mixed
    with a uniform pseudocount of the given weight.  Returned as q x q with the last state's
    row and column 0.  Computed by a full matrix inverse."""
    n = q - 1
    f = np.asarray(N, dtype=np.float64) / B
    lam = pseudocount
    fij = (1 - lam) * f + lam / (q * q)
    fr = (1 - lam) * f.sum(1) + lam / q
    fj = (1 - lam) * f.sum(0) + lam / q
    C = np.zeros((2 * n, 2 * n))
    C[:n, :n] = np.diag(fr[:n]) - np.outer(fr[:n], fr[:n])
    C[n:, n:] = np.diag(fj[:n]) - np.outer(fj[:n], fj[:n])
    C[:n, n:] = fij[:n, :n] - np.outer(fr[:n], fj[:n])
    C[n:, :n] = C[:n, n:].T
    out = np.zeros((q, q))
    out[:n, :n] = -np.linalg.inv(C)[:n, n:]
    return out


def pair_features(N, B: int, q: int) -> np.ndarray:
    """Features of one pair from its counts: joint frequencies [q*q], the marginals of r and
    of j [q each] and the mean-field block [q*q]."""
    f = np.asarray(N, dtype=np.float64) / B
    return np.concatenate([f.ravel(), f.sum(1), f.sum(0), mean_field_block(N, B, q).ravel()])
