"""Pair statistics on the GPU (no SuperDCA counterpart): exact counts, MI, top-K partners.

Counts are pinned bit-exact by numpy's integer matrix product of the one-hot indicators;
mutual information by scikit-learn's mutual_info_score on the label vectors; the top-K
lists by a full sort of the FP64 MI rows.
"""
import numpy as np
import pytest
from sklearn.metrics import mutual_info_score

import superdca_amd as sd
from superdca_amd import synth

pytestmark = pytest.mark.gpu


def onehot(states, q, QP):
    B, L = states.shape
    X = np.zeros((B, L * QP), dtype=np.int64)
    for a in range(q):
        X[:, a::QP] = states == a
    return X


def mi_matrix(states, q):
    B, L = states.shape
    X = onehot(states, q, q).astype(np.float64)
    N = X.T @ X
    f = N / B
    fa = np.diag(f).reshape(L, q)
    MI = np.zeros((L, L))
    F = f.reshape(L, q, L, q)
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = F / (fa[:, :, None, None] * fa[None, None, :, :])
        t = np.where(F > 0, F * np.log(ratio), 0.0)
    MI = t.sum(axis=(1, 3))
    np.fill_diagonal(MI, -1.0)
    return MI


def test_count_table_bit_exact_int8_mfma(c1):
    st = c1["states"]
    X = onehot(st, 4, 4)
    with sd.Context(st) as ctx:
        for i0, n in ((0, 64), (937, 63), (500, 1)):
            tab = ctx.pair_count_table(i0, n)
            ref = X[:, i0 * 4:(i0 + n) * 4].T @ X
            np.testing.assert_array_equal(tab, ref)


def test_count_table_bit_exact_q5():
    d = synth.generate(B=333, N=150, clades=7, pairs=3, seed=8)
    st, _ = sd.prepare_alignment(d["chars"], q=5)
    X = onehot(st, 5, 8)
    with sd.Context(st, q=5) as ctx:
        tab = ctx.pair_count_table(10, 40)
    np.testing.assert_array_equal(tab, X[:, 80:400].T @ X)


def test_pair_counts_bit_exact(c1):
    st = c1["states"]
    rng = np.random.default_rng(3)
    pairs = rng.integers(0, st.shape[1], (500, 2)).astype(np.int32)
    pairs[:5, 1] = pairs[:5, 0]  # include i == j
    with sd.Context(st) as ctx:
        cnt = ctx.pair_counts(pairs)
    for p, (i, j) in enumerate(pairs):
        ref = np.zeros((4, 4), dtype=np.int64)
        np.add.at(ref, (st[:, i], st[:, j]), 1)
        np.testing.assert_array_equal(cnt[p], ref)


def test_mi_matches_sklearn_and_topk_is_exact(c1):
    st = c1["states"]
    L = st.shape[1]
    K = 64
    with sd.Context(st) as ctx:
        idx, val = ctx.mi_topk(K)
    MI = mi_matrix(st, 4)
    # MI of the selected partners against scikit-learn for a sample of loci
    for i in (0, 17, 503, 999):
        for k in (0, 31, 63):
            j = idx[i, k]
            assert val[i, k] == pytest.approx(mutual_info_score(st[:, i], st[:, j]), rel=2e-5, abs=1e-7)
    overlap = 0
    for i in range(L):
        assert i not in idx[i]
        assert np.all(np.diff(idx[i]) > 0)  # ascending index order
        ref = np.argsort(-MI[i], kind="stable")[:K]
        # every chosen partner must be within FP32 rounding of the K-th best value
        kth = MI[i, ref[-1]]
        assert np.all(MI[i, idx[i]] >= kth - 1e-6)
        overlap += len(set(ref) & set(idx[i]))
    assert overlap / (L * K) >= 0.999


def test_topk_q5():
    d = synth.generate(B=400, N=300, clades=8, pairs=6, seed=12)
    st, loci = sd.prepare_alignment(d["chars"], q=5)
    assert loci.size == 300
    with sd.Context(st, q=5) as ctx:
        idx, val = ctx.mi_topk(10)
    MI = mi_matrix(st, 5)
    for i in range(0, 300, 37):
        ref = set(np.argsort(-MI[i], kind="stable")[:10])
        assert len(ref & set(idx[i])) >= 9
