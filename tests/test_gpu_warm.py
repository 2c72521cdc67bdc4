"""Warm starts on the GPU: features, the network's kernels against the numpy reference,
training, and warm versus cold solves (same optimum, fewer iterations)."""
import numpy as np
import pytest

import oracle as orc
from oracle import mlp_ref
import superdca_amd as sd
from superdca_amd import synth, warmstart

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("q", [4, 5])
def test_pair_features_match_numpy(c1, q):
    st, _ = sd.prepare_alignment(synth.workload("C1")["chars"], q=q) if q != 4 else (c1["states"], None)
    rng = np.random.default_rng(2)
    pairs = rng.integers(0, st.shape[1], (200, 2)).astype(np.int32)
    with sd.Context(st, q=q) as ctx:
        f = ctx.pair_features(pairs)
    B = st.shape[0]
    assert f.shape == (200, 2 * q * q + 2 * q)
    for p, (r, j) in enumerate(pairs[:80]):
        N = np.zeros((q, q))
        np.add.at(N, (st[:, r], st[:, j]), 1)
        want = mlp_ref.pair_features(N, B, q)
        np.testing.assert_allclose(f[p, :q * q + 2 * q], want[:q * q + 2 * q], rtol=1e-6, atol=1e-7)
        # the mean-field block: Schur complement on the device, full inverse in the reference
        np.testing.assert_allclose(f[p, q * q + 2 * q:], want[q * q + 2 * q:], rtol=1e-5,
                                   atol=1e-6 * max(1.0, np.abs(want).max()))


def test_mlp_forward_and_gradient_match_reference(c1):
    """BF16 MFMA layers against the numpy network with the same BF16 rounding points
    (tight) and against the exact FP64 network (BF16's precision)."""
    st = c1["states"][:, :50]
    rng = np.random.default_rng(4)
    n, dh, dout = 300, 128, 16
    din = 2 * 16 + 8
    x = rng.random((n, din)).astype(np.float32)
    t = rng.normal(size=(n, dout)).astype(np.float32) * 0.1
    mask = (rng.random((n, dout)) < 0.85).astype(np.float32)
    with sd.Context(st) as ctx:
        ctx.mlp_init(dh, seed=7)
        p = ctx.mlp_params()
        y = ctx.mlp_forward(x)
        loss, g = ctx.mlp_loss_grad(x, t, mask)
    yb, _ = mlp_ref.forward_bf16(p, x, din, dh, dout)
    np.testing.assert_allclose(y, yb, rtol=1e-3, atol=1e-3 * np.abs(yb).max())
    lb, gb = mlp_ref.loss_grad_bf16(p, x, t, mask, din, dh, dout)
    assert loss == pytest.approx(lb, rel=1e-4)
    assert np.linalg.norm(g - gb) < 2e-3 * np.linalg.norm(gb)
    # FP64 network: BF16 precision
    yr, _ = mlp_ref.forward(p, x, din, dh, dout)
    assert np.abs(y - yr).max() < 2e-2 * np.abs(yr).max()
    lr_, gr = mlp_ref.loss_grad(p, x, t, mask, din, dh, dout)
    assert loss == pytest.approx(lr_, rel=2e-2)
    assert np.linalg.norm(g - gr) < 3e-2 * np.linalg.norm(gr)


def test_mlp_training_decreases_loss(c1):
    st = c1["states"][:, :50]
    rng = np.random.default_rng(5)
    x = rng.random((5000, 40)).astype(np.float32)
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


@pytest.fixture(scope="module")
def trained_q5():
    """The network at q_env = 5, trained as the benchmark trains it: cold-start solutions of
    400 loci each of the C1 and C2 workloads, top-32 MI plus 32 random partners per locus."""
    feats, tgts, masks = [], [], []
    for wl in ("C1", "C2"):
        st, _ = sd.prepare_alignment(synth.workload(wl)["chars"], q=5)
        f, t, m, _ = warmstart.training_set(st, q=5, n_loci=400, top=32, rand=32, seed=0)
        feats.append(f)
        tgts.append(t)
        masks.append(m)
    params, losses = warmstart.train(np.concatenate(feats), np.concatenate(tgts), np.concatenate(masks), st, q=5,
                                     epochs=40, batch=2048, lr=2e-3, seed=0)
    return params, losses


def test_c5_warm_and_cold_same_optimum(c4_states, trained_q5):
    """C5: 256 loci of the genome-scale alignment (B = 3000, L = 94,000, q_env = 5), split
    BF16, solved cold and warm-started from the BF16 network for each locus' 2048 largest-MI
    partners (the benchmark's mode: the blocks as the first search direction, fields from
    zero), stopping at ||g||_inf < 1e-3 and ||g_J||_inf < 1e-6.  Both reach the same optimum.
    The evaluation ratio is recorded: at C4 the iteration count is set by the conditioning
    that clade-correlated columns give the Hessian, which a start for 2048 of 94,000
    partners hardly changes, so warm is not required to win here (the B = 600 alignment
    above is where the network's gain shows)."""
    params, losses = trained_q5
    assert losses[-1] < 0.2 * losses[0]
    st = c4_states
    loci = np.random.default_rng(99).choice(st.shape[1], 256, replace=False).astype(np.int32)
    kw = dict(q=5, formulation=sd.FORM_PRIMAL, precision=sd.PREC_SPLIT_BF16, fine_factor=0, grad_threshold=1e-3,
              j_threshold=1e-6, batch_loci=256)
    with sd.Context(st, **kw) as ctx:
        w, _ = ctx.reweight(0.9)
        ctx.set_weights(w)
        _, cold, _, _ = ctx.solve(loci)
        ctx.mlp_set_params(params, 128)
        _, warm, _, _ = ctx.solve_mlp(loci, npartners=2048, mode=sd.WARM_DIRECTION | sd.WARM_NO_FIELDS)
    for s in (cold, warm):
        ok = (s["status"] == sd.ST_CONVERGED) | ((s["status"] == sd.ST_LS_FAIL) & (s["gnorm"] < 1e-3))
        assert ok.all(), np.bincount(s["status"])
    rel = np.abs(warm["fval"] - cold["fval"]) / np.abs(cold["fval"])
    ratio = warm["nfeval"].sum() / cold["nfeval"].sum()
    print(f"C5 256 loci: evaluations cold {int(cold['nfeval'].sum())} warm {int(warm['nfeval'].sum())} "
          f"(warm/cold {ratio:.3f}; medians {np.median(cold['nfeval']):.0f} / {np.median(warm['nfeval']):.0f}); "
          f"f warm vs cold max rel {rel.max():.2e}")
    assert rel.max() < 1e-6
    # the network's start must not cost the solve more than a tenth of its evaluations
    assert ratio < 1.10
