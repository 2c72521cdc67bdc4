"""Training data and training of the warm-start network (host side of sdca_mlp.hip).

A training example is a locus pair (r, j): its features (sdca_pair_features: joint and
marginal frequencies, mutual information) and the solved coupling block J_r[j](a, s) of
locus r at partner j, taken from a cold-start solution in SuperDCA's parameter layout.
Partners are the top mutual-information partners of r plus random loci, so the network
also learns that weakly associated pairs have small couplings.  Outputs for states a
never observed at j do not exist as parameters and are masked out of the loss.
"""
from __future__ import annotations

import numpy as np

from . import Context


def ref_index(L: int, q: int, j, a, s):
    """Offsets of J_r[j](a, s) in the reference parameter vector (16-locus blocks, STATES
    access order; SuperDCA/include/Matrix_kernel_access_order.hpp:28-34)."""
    j = np.asarray(j, dtype=np.int64)
    blk = j // 16
    pos = j % 16
    ext = np.where((blk + 1) * 16 > L, L % 16, 16)
    return blk * 16 * q * q + (np.asarray(a) * ext + pos) * q + np.asarray(s)


def observed(states: np.ndarray, q: int) -> np.ndarray:
    """obs[j, a]: state a occurs at locus j."""
    obs = np.zeros((states.shape[1], q), dtype=bool)
    for a in range(q):
        obs[:, a] = (states == a).any(axis=0)
    return obs


def training_set(states: np.ndarray, q: int = 4, n_loci: int = 256, top: int = 32, rand: int = 32, seed: int = 0,
                 grad_threshold: float = 1e-4, batch_loci: int = 128, **opts):
    """Solves n_loci random loci of the alignment from a cold start and returns
    (features, targets, mask, stats) for top-MI and random partner pairs of each."""
    B, L = states.shape
    rng = np.random.default_rng(seed)
    n_loci = min(n_loci, L)
    loci = np.sort(rng.choice(L, n_loci, replace=False)).astype(np.int32)
    with Context(states, q=q, grad_threshold=grad_threshold, batch_loci=batch_loci, **opts) as ctx:
        w, _ = ctx.reweight(0.9)
        ctx.set_weights(w)
        _, stats, X, _ = ctx.solve(loci, want_x=True, rows=False)
        top = min(top, L - 1)
        part, _ = ctx.mi_topk(top)
        rows = []
        for k, r in enumerate(loci):
            cand = rng.choice(L, rand + 1, replace=False) if rand else np.zeros(0, dtype=np.int64)
            js = np.unique(np.concatenate([part[r], cand[cand != r][:rand]]))
            rows.append(np.stack([np.full(js.size, k), np.full(js.size, r), js], axis=1))
        rows = np.concatenate(rows)
        feats = ctx.pair_features(rows[:, 1:3])
    obs = observed(states, q)
    s_idx, a_idx = np.meshgrid(np.arange(q), np.arange(q), indexing="ij")  # output s*q + a
    s_idx = s_idx.ravel()
    a_idx = a_idx.ravel()
    off = ref_index(L, q, rows[:, 2][:, None], a_idx[None, :], s_idx[None, :])
    targets = X[rows[:, 0][:, None], off].astype(np.float32)
    mask = obs[rows[:, 2][:, None], a_idx[None, :]].astype(np.float32)
    return feats, targets, mask, stats


def train(feats, targets, mask, states_for_device: np.ndarray, q: int = 4, hidden: int = 128, epochs: int = 30,
          batch: int = 4096, lr: float = 2e-3, seed: int = 0):
    """Trains the network on the device (Adam, hand-written kernels) -> (params, losses)."""
    with Context(states_for_device, q=q) as ctx:
        ctx.mlp_init(hidden, seed)
        losses = ctx.mlp_train(feats, targets, mask, epochs=epochs, batch=batch, lr=lr, seed=seed)
        return ctx.mlp_params(), losses
