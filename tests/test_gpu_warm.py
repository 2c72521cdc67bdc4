"""Warm starts on the GPU: features, the network's kernels against the numpy reference,
training, and warm versus cold solves (same optimum, fewer iterations)."""
import numpy as np
import pytest

import oracle as orc
from oracle import mlp_ref
import superdca_amd as sd
from superdca_amd import synth, warmstart

pytestmark = pytest.mark.gpu


def test_pair_features_match_numpy(c1):
    st = c1["states"]
    rng = np.random.default_rng(2)
    pairs = rng.integers(0, st.shape[1], (200, 2)).astype(np.int32)
    with sd.Context(st) as ctx:
        f = ctx.pair_features(pairs)
    B, q = st.shape[0], 4
    for p, (r, j) in enumerate(pairs[:50]):
        N = np.zeros((q, q))
        np.add.at(N, (st[:, r], st[:, j]), 1)
        F = N / B
        np.testing.assert_allclose(f[p, :16], F.ravel(), rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(f[p, 16:20], F.sum(1), rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(f[p, 20:24], F.sum(0), rtol=1e-6, atol=1e-7)
        from sklearn.metrics import mutual_info_score
        assert f[p, 24] == pytest.approx(mutual_info_score(st[:, r], st[:, j]), rel=1e-4, abs=1e-6)


def test_mlp_forward_and_gradient_match_reference(c1):
    st = c1["states"][:, :50]
    rng = np.random.default_rng(4)
    n, din, dh, dout = 300, 25, 128, 16
    x = rng.random((n, din)).astype(np.float32)
    t = rng.normal(size=(n, dout)).astype(np.float32) * 0.1
    mask = (rng.random((n, dout)) < 0.85).astype(np.float32)
    with sd.Context(st) as ctx:
        ctx.mlp_init(dh, seed=7)
        p = ctx.mlp_params()
        y = ctx.mlp_forward(x)
        loss, g = ctx.mlp_loss_grad(x, t, mask)
    yr, _ = mlp_ref.forward(p, x, din, dh, dout)
    np.testing.assert_allclose(y, yr, rtol=1e-4, atol=1e-5)
    lr_, gr = mlp_ref.loss_grad(p, x, t, mask, din, dh, dout)
    assert loss == pytest.approx(lr_, rel=1e-5)
    np.testing.assert_allclose(g, gr, rtol=1e-3, atol=1e-6 * np.abs(gr).max())


def test_mlp_training_decreases_loss(c1):
    st = c1["states"][:, :50]
    rng = np.random.default_rng(5)
    x = rng.random((5000, 25)).astype(np.float32)
    t = np.stack([np.sin(3 * x[:, k]) * x[:, k + 1] for k in range(16)], axis=1).astype(np.float32)
    with sd.Context(st) as ctx:
        ctx.mlp_init(128, seed=1)
        losses = ctx.mlp_train(x, t, epochs=20, batch=512, lr=2e-3, seed=3)
        p1 = ctx.mlp_params()
        ctx.mlp_init(128, seed=1)
        losses2 = ctx.mlp_train(x, t, epochs=20, batch=512, lr=2e-3, seed=3)
        p2 = ctx.mlp_params()
    assert losses[-1] < 0.35 * losses[0]
    np.testing.assert_array_equal(p1, p2)  # deterministic
    np.testing.assert_array_equal(losses, losses2)


@pytest.fixture(scope="module")
def trained_c1(c1):
    feats, tgt, mask, _ = warmstart.training_set(c1["states"], n_loci=400, top=32, rand=32, seed=0)
    params, losses = warmstart.train(feats, tgt, mask, c1["states"], epochs=40, batch=2048, lr=2e-3)
    return params, losses


def test_warm_start_same_optimum_fewer_iterations(trained_c1):
    params, losses = trained_c1
    assert losses[-1] < 0.05 * losses[0]
    d = synth.generate(B=600, N=1500, clades=20, pairs=60, seed=4242)
    st, _ = sd.prepare_alignment(d["chars"])
    loci = np.arange(0, st.shape[1], 23, dtype=np.int32)
    with sd.Context(st, grad_threshold=1e-4, batch_loci=64) as ctx:
        w, _ = ctx.reweight(0.9)
        ctx.set_weights(w)
        _, sc, _, rc = ctx.solve(loci)
        ctx.mlp_set_params(params, 128)
        runs = {"cold": sc}
        for name, mode in (("point", sd.WARM_POINT | sd.WARM_NO_FIELDS), ("point_fields", sd.WARM_POINT),
                           ("direction", sd.WARM_DIRECTION | sd.WARM_NO_FIELDS)):
            runs[name] = ctx.solve_mlp(loci, npartners=512, mode=mode)[1]
        ctx.mlp_set_params(np.zeros_like(params), 128)  # network off: the J = 0 field optimum only
        runs["fields_only"] = ctx.solve_mlp(loci, npartners=512, mode=sd.WARM_POINT)[1]
    for s in runs.values():
        ok = (s["status"] == sd.ST_CONVERGED) | ((s["status"] == sd.ST_LS_FAIL) & (s["gnorm"] < 3e-4))
        assert ok.all()
        # strictly convex: every start reaches the same optimum
        np.testing.assert_allclose(s["fval"], sc["fval"], rtol=1e-6)
    ev = {k: int(s["nfeval"].sum()) for k, s in runs.items()}
    print(f"evaluations {ev}")
    # what the trained network gains over the method without it: the network's blocks as
    # the start point (fields at zero) against the cold start
    assert ev["point"] < 0.95 * ev["cold"]
