"""Pins the FP64 CPU oracle against library routines, closed forms and invariants.

Each check is chosen so that a plausible slip in the restatement (a dropped term, a wrong
sign or index, a transposed operand) fails at least one of them.
"""
import numpy as np
import pytest
from scipy.optimize import minimize
from scipy.spatial.distance import pdist

import oracle as orc

GOLD = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


def rand_states(rng, B, L, q=4, p_gap=0.1):
    st = rng.integers(0, q - 1, (B, L)).astype(np.uint8)
    st[rng.random((B, L)) < p_gap] = q - 1
    return st


# ----------------------------------------------------------------- layout

@pytest.mark.parametrize("L,q", [(1, 4), (15, 4), (16, 4), (17, 4), (40, 4), (33, 5)])
def test_ref_index_is_bijection(L, q):
    idx = sorted(orc.ref_index(L, q, j, a, s) for j in range(L) for a in range(q) for s in range(q))
    assert idx == list(range(L * q * q))
    assert orc.nparams(L, q) == L * q * q + q


def test_ref_index_known_positions():
    # L = 20: block 0 has 16 loci, block 1 has 4 (extent L % 16).  STATES order:
    # (state * extent + position) * q + s inside a block of 16*q*q values.
    q = 4
    assert orc.ref_index(20, q, 0, 0, 0) == 0
    assert orc.ref_index(20, q, 0, 0, 3) == 3
    assert orc.ref_index(20, q, 1, 0, 0) == 4          # next position, same state
    assert orc.ref_index(20, q, 0, 1, 0) == 16 * 4      # next state row starts after 16 positions
    assert orc.ref_index(20, q, 16, 0, 0) == 256        # block 1 base = 16*q*q
    assert orc.ref_index(20, q, 17, 2, 1) == 256 + (2 * 4 + 1) * 4 + 1


# ------------------------------------------------------------ reweighting

def test_hamming_matches_scipy_pdist():
    rng = np.random.default_rng(1)
    st = rand_states(rng, 37, 101)
    d = orc.hamming(st)
    ref = np.rint(pdist(st, metric="hamming") * st.shape[1]).astype(np.int64)
    # pdist order is (0,1),(0,2),...; the oracle packs i*(i-1)/2 + j for j < i
    B = st.shape[0]
    mine = np.array([d[i * (i - 1) // 2 + j] for j in range(B) for i in range(j + 1, B)])
    np.testing.assert_array_equal(mine, ref)


def test_weights_special_cases():
    rng = np.random.default_rng(2)
    st = rand_states(rng, 20, 50)
    w, iw = orc.weights(st, theta=1.5)  # nobody is that similar
    np.testing.assert_array_equal(w, np.ones(20))
    w, iw = orc.weights(st, theta=0.0)  # everybody is a neighbour
    np.testing.assert_allclose(w, np.full(20, 1 / 20), rtol=0, atol=0)
    same = np.repeat(st[:1], 7, axis=0)
    w, _ = orc.weights(same, theta=0.9)
    np.testing.assert_array_equal(w, np.full(7, 1 / 7))
    # duplicated pairs and otherwise far apart sequences: weight 1/2 each
    base = np.zeros((3, 60), dtype=np.uint8)
    base[1, :] = 1
    base[2, :] = 2
    dup = np.concatenate([base, base])
    w, iw = orc.weights(dup, theta=0.9)
    np.testing.assert_array_equal(w, np.full(6, 0.5))
    np.testing.assert_array_equal(iw, np.full(6, 2))


def test_weights_threshold_is_inclusive_and_uses_multiplicity():
    # identity exactly 0.9 (9 of 10 positions equal) counts as a neighbour (>=, :72)
    a = np.zeros((2, 10), dtype=np.uint8)
    a[1, 0] = 1
    w, iw = orc.weights(a, theta=0.9)
    np.testing.assert_array_equal(iw, [2, 2])
    w, iw = orc.weights(a, theta=0.9000001)
    np.testing.assert_array_equal(iw, [1, 1])
    w, iw = orc.weights(a, theta=0.9, mult=np.array([3, 5], dtype=np.uint32))
    np.testing.assert_array_equal(iw, [8, 8])
    np.testing.assert_allclose(w, [3 / 8, 5 / 8], rtol=0, atol=0)


def test_lambda_closed_form():
    w = np.full(100, 0.5)  # B_eff = 50 <= 500 -> auto = 1 - 0.9*50/500 = 0.91
    beff, lh, lJ = orc.lambdas(w)
    assert beff == 50.0
    assert lh == pytest.approx(0.91 * 50, rel=1e-15)
    assert lJ == pytest.approx(0.91 * 50 / 2, rel=1e-15)
    w = np.ones(1000)  # B_eff > 500 -> 0.1
    beff, lh, lJ = orc.lambdas(w)
    assert (lh, lJ) == (pytest.approx(100.0), pytest.approx(50.0))
    beff, lh, lJ = orc.lambdas(w, lambda_h=0.3, lambda_J=0.02)
    assert (lh, lJ) == (pytest.approx(300.0), pytest.approx(20.0))


# -------------------------------------------------------------- objective

def setup_small(seed=3, B=50, L=21):
    rng = np.random.default_rng(seed)
    st = rand_states(rng, B, L)
    w = rng.uniform(0.1, 1.0, B)
    x = rng.normal(size=orc.nparams(L, 4)) * 0.3
    return rng, st, w, x


def test_objective_finite_differences():
    rng, st, w, x = setup_small()
    r = 17  # in the short last block
    f, g = orc.objective(st, r, w, 2.0, 1.5, x)
    h = 1e-6
    idx = rng.choice(x.size, 40, replace=False)
    for i in idx:
        e = np.zeros_like(x)
        e[i] = h
        fd = (orc.objective(st, r, w, 2.0, 1.5, x + e)[0] - orc.objective(st, r, w, 2.0, 1.5, x - e)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i]))


@pytest.mark.parametrize("q", [3, 5])
def test_objective_finite_differences_other_envelopes(q):
    # the softmax spans the whole q-state envelope, unobserved states included
    rng = np.random.default_rng(30 + q)
    B, L, r = 40, 19, 16
    st = rand_states(rng, B, L, q=q)
    w = rng.uniform(0.1, 1.0, B)
    x = rng.normal(size=orc.nparams(L, q)) * 0.3
    f, g = orc.objective(st, r, w, 2.0, 1.5, x, q=q)
    h = 1e-6
    for i in rng.choice(x.size, 40, replace=False):
        e = np.zeros_like(x)
        e[i] = h
        fd = (orc.objective(st, r, w, 2.0, 1.5, x + e, q=q)[0] - orc.objective(st, r, w, 2.0, 1.5, x - e, q=q)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i]))
    own = [orc.ref_index(L, q, r, a, s) for a in range(q) for s in range(q)]
    assert np.all(g[own] == 0.0)


def test_objective_excludes_own_block():
    rng, st, w, x = setup_small()
    r, L, q = 5, st.shape[1], 4
    f, g = orc.objective(st, r, w, 2.0, 1.5, x)
    x2 = x.copy()
    for a in range(q):
        for s in range(q):
            x2[orc.ref_index(L, q, r, a, s)] += 10.0
    f2, g2 = orc.objective(st, r, w, 2.0, 1.5, x2)
    assert f2 == f
    own = [orc.ref_index(L, q, r, a, s) for a in range(q) for s in range(q)]
    assert np.all(g[own] == 0.0)


def test_objective_matches_sklearn_logistic_regression():
    d = np.load(f"{GOLD}/logreg_L2.npz")
    st, w, lam = d["states"], d["w"], float(d["lam"])
    q, L = 4, 2
    x = np.zeros(orc.nparams(L, q))
    for a in range(q):
        for s in range(q):
            x[orc.ref_index(L, q, 1, a, s)] = d["J1"][a, s]
    x[L * q * q:] = d["h"]
    f, g = orc.objective(st, 0, w, lam, lam, x)
    # scikit-learn's optimum is a stationary point of the oracle's objective
    assert np.abs(g).max() < 1e-6 * max(1.0, f)
    # and the oracle's own solver lands on the same parameters
    _, X, stats = orc.solve_loci(st, w, lam, lam, [0], q=q, gtol=1e-8, max_iter=5000, nthreads=1,
                                 blocked=False, want_rows=False, want_x=True)
    assert stats["status"][0] == 0
    np.testing.assert_allclose(X[0], x, atol=2e-6)


def test_objective_h_closed_form_when_J_zero():
    # with every J = 0 the h-gradient vanishes at h = argmin W*lse(h) - c.h + lh*|h|^2
    rng, st, w, _ = setup_small(seed=4)
    r, L, q, lh = 7, st.shape[1], 4, 3.0
    c = np.array([w[st[:, r] == s].sum() for s in range(q)])
    W = w.sum()
    h = np.zeros(q)
    for _ in range(50):  # Newton on the q-dimensional strictly convex problem
        p = np.exp(h - h.max())
        p /= p.sum()
        grad = W * p - c + 2 * lh * h
        H = W * (np.diag(p) - np.outer(p, p)) + 2 * lh * np.eye(q)
        h -= np.linalg.solve(H, grad)
    x = np.zeros(orc.nparams(L, q))
    x[L * q * q:] = h
    f, g = orc.objective(st, r, w, lh, 1.0, x)
    np.testing.assert_allclose(g[L * q * q:], 0.0, atol=1e-10)
    lse = np.log(np.exp(h).sum())
    assert f == pytest.approx(W * lse - c @ h + lh * h @ h, rel=1e-13)


def test_objective_sequence_permutation_and_unit_weights():
    rng, st, w, x = setup_small(seed=5)
    r = 3
    f, g = orc.objective(st, r, w, 2.0, 1.5, x)
    perm = rng.permutation(st.shape[0])
    f2, g2 = orc.objective(st[perm], r, w[perm], 2.0, 1.5, x)
    assert f2 == pytest.approx(f, rel=1e-13)
    np.testing.assert_allclose(g2, g, rtol=1e-12, atol=1e-12)
    # doubling every weight doubles the data term
    f0, _ = orc.objective(st, r, w, 0.0, 0.0, x)
    f1, _ = orc.objective(st, r, 2 * w, 0.0, 0.0, x)
    assert f1 == pytest.approx(2 * f0, rel=1e-13)


@pytest.mark.parametrize("r", [0, 15, 16, 40, 44])
def test_blocked_objective_equals_plain(r):
    rng = np.random.default_rng(6)
    B, L = 64, 45  # last block holds 13 loci
    base = rand_states(rng, 4, L)
    st = base[rng.integers(0, 4, B)]  # few unique patterns per block
    flip = rng.random((B, L)) < 0.05
    st[flip] = rng.integers(0, 4, flip.sum())
    w = rng.uniform(0.1, 1.0, B)
    x = rng.normal(size=orc.nparams(L, 4)) * 0.2
    f, g = orc.objective(st, r, w, 1.3, 0.7, x)
    blk = orc.Blocks(st)
    assert blk.npatterns < B * ((L + 15) // 16)
    fb, gb = blk.objective(r, w, 1.3, 0.7, x)
    assert fb == pytest.approx(f, rel=1e-13)
    np.testing.assert_allclose(gb, g, rtol=1e-11, atol=1e-12)


# ---------------------------------------------------------------- L-BFGS

def test_lbfgs_reaches_scipy_optimum(c1):
    st = c1["states"][:, :120]
    w, _ = orc.weights(st)
    beff, lh, lJ = orc.lambdas(w)
    r = 31
    _, X, stats = orc.solve_loci(st, w, lh, lJ, [r], gtol=1e-7, max_iter=5000, nthreads=1, want_rows=False,
                                 want_x=True)
    assert stats["status"][0] == 0
    f_orc = stats["fval"][0]
    res = minimize(lambda v: orc.objective(st, r, w, lh, lJ, v), np.zeros(orc.nparams(st.shape[1], 4)), jac=True,
                   method="L-BFGS-B", options=dict(maxiter=20000, gtol=1e-10, ftol=1e-15, maxcor=20))
    assert f_orc == pytest.approx(res.fun, rel=1e-10)
    # strong-convexity certificate: |x - x*| <= |g| / (2 min(lh, lJ))
    _, gsc = orc.objective(st, r, w, lh, lJ, res.x)
    bound = (np.linalg.norm(gsc) + np.sqrt(X.shape[1]) * 1e-7) / (2 * min(lh, lJ))
    assert np.linalg.norm(X[0] - res.x) <= bound + 1e-7


def test_lbfgs_reference_mode_stops_at_threshold(c1):
    st = c1["states"][:, :200]
    w, _ = orc.weights(st)
    beff, lh, lJ = orc.lambdas(w)
    _, _, stats = orc.solve_loci(st, w, lh, lJ, np.arange(8), gtol=1e-3, nthreads=4, want_rows=False)
    assert np.all(stats["status"] == 0)
    assert np.all(stats["gnorm"] < 1e-3)
    assert np.all(stats["nfeval"] >= stats["iter"] + 1)


def test_warm_and_cold_start_same_optimum(c1):
    st = c1["states"][:, :100]
    w, _ = orc.weights(st)
    beff, lh, lJ = orc.lambdas(w)
    loci = [3, 50]
    _, Xc, sc = orc.solve_loci(st, w, lh, lJ, loci, gtol=1e-7, max_iter=5000, want_rows=False, want_x=True)
    rng = np.random.default_rng(9)
    x0 = Xc + rng.normal(size=Xc.shape) * 0.05
    _, Xw, sw = orc.solve_loci(st, w, lh, lJ, loci, gtol=1e-7, max_iter=5000, x0=x0, want_rows=False, want_x=True)
    np.testing.assert_allclose(sw["fval"], sc["fval"], rtol=1e-12)


# ---------------------------------------------------------- gauge + score

def test_gauge_hand_example():
    # M with row means r = [1.5, 5.5, 9.5, 13.5] and column means c = [6, 7, 8, 9],
    # grand mean g = 7.5.  The reference gauge is M[i][k] - r[k] - c[i] + g.
    M = np.arange(16, dtype=float).reshape(4, 4)
    G = orc.ising_gauge(M)
    r = M.mean(axis=1)
    c = M.mean(axis=0)
    expected = M - r[None, :] - c[:, None] + 7.5
    np.testing.assert_array_equal(G, expected)
    assert G[0, 0] == 0 - 1.5 - 6 + 7.5
    assert G[3, 1] == 13 - 5.5 - 9 + 7.5


@pytest.mark.parametrize("q", [3, 4, 5])
def test_gauge_symmetric_equals_textbook_and_invariances(q):
    rng = np.random.default_rng(10)
    A = rng.normal(size=(q, q))
    S = A + A.T
    textbook = S - S.mean(axis=1, keepdims=True) - S.mean(axis=0, keepdims=True) + S.mean()
    np.testing.assert_allclose(orc.ising_gauge(S), textbook, atol=1e-14)
    # the reference variant differs from the textbook one for non-symmetric blocks
    tb = A - A.mean(axis=1, keepdims=True) - A.mean(axis=0, keepdims=True) + A.mean()
    assert np.abs(orc.ising_gauge(A) - tb).max() > 1e-3
    # transposed input gives the transposed gauge; adding a constant changes nothing
    np.testing.assert_allclose(orc.ising_gauge(A, transpose=True), orc.ising_gauge(A).T, atol=1e-14)
    np.testing.assert_allclose(orc.ising_gauge(A + 3.25), orc.ising_gauge(A), atol=1e-13)


def test_frobenius_excludes_gap_row_and_column():
    G = np.arange(16, dtype=float).reshape(4, 4)
    assert orc.frobenius_excl(G, 3) == pytest.approx(np.sqrt((G[:3, :3] ** 2).sum()), rel=1e-15)
    assert orc.frobenius_excl(G, 0) == pytest.approx(np.sqrt((G[1:, 1:] ** 2).sum()), rel=1e-15)
    keep = [0, 2, 3]
    assert orc.frobenius_excl(G, 1) == pytest.approx(np.sqrt((G[np.ix_(keep, keep)] ** 2).sum()), rel=1e-15)


@pytest.mark.parametrize("q", [3, 4, 5])
def test_score_row_reads_block_of_each_partner(q):
    # every envelope width the engine uses (gap = state q-1)
    L, r = 20, 6
    rng = np.random.default_rng(11)
    x = rng.normal(size=orc.nparams(L, q))
    row = orc.score_row(x, L, q, r)
    assert row[r] == 0.0
    for n in (0, 5, 7, 16, 19):
        M = np.array([[x[orc.ref_index(L, q, n, a, s)] for s in range(q)] for a in range(q)])
        G = orc.ising_gauge(M, transpose=n < r)
        assert row[n] == np.float32(np.sqrt((G[:q - 1, :q - 1] ** 2).sum()))


# ------------------------------------------------------- warm-start network

def test_mlp_reference_gradient_finite_differences():
    from oracle import mlp_ref

    rng = np.random.default_rng(13)
    din, dh, dout, n = 25, 12, 16, 9
    npar = dh * din + dh + dh * dh + dh + dout * dh + dout
    p = rng.normal(size=npar) * 0.4
    x = rng.random((n, din))
    t = rng.normal(size=(n, dout))
    mask = (rng.random((n, dout)) < 0.8).astype(float)
    loss, g = mlp_ref.loss_grad(p, x, t, mask, din, dh, dout)
    h = 1e-6
    for i in rng.choice(npar, 60, replace=False):
        e = np.zeros(npar)
        e[i] = h
        fd = (mlp_ref.loss_grad(p + e, x, t, mask, din, dh, dout)[0] -
              mlp_ref.loss_grad(p - e, x, t, mask, din, dh, dout)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i]))
    # the loss is the masked mean squared error
    y, _ = mlp_ref.forward(p, x, din, dh, dout)
    assert loss == pytest.approx((((y - t) * mask) ** 2).sum() / mask.sum(), rel=1e-14)


def test_bf16_rounding_matches_torch():
    import torch
    from oracle import mlp_ref

    rng = np.random.default_rng(21)
    x = np.concatenate([rng.normal(size=4000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 4000),
                        np.array([0.0, -0.0, 1.0, np.inf, -np.inf, 3.4e38, 1e-40], dtype=np.float32)])
    # halfway cases: low 16 bits exactly 0x8000 round to the even neighbour
    half = (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    x = np.concatenate([x, half, -half]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(mlp_ref.bf16(x), want)
    assert np.isnan(mlp_ref.bf16(np.array([np.nan], dtype=np.float32))[0])


def test_mlp_bf16_gradient_close_to_fp64():
    """The BF16 rounding points change the loss and gradient by BF16's relative precision,
    not more: a dropped rounding point or a transposed operand would not stay this close."""
    from oracle import mlp_ref

    rng = np.random.default_rng(17)
    din, dh, dout, n = 40, 64, 16, 300
    npar = dh * din + dh + dh * dh + dh + dout * dh + dout
    p = rng.normal(size=npar) * 0.15
    x = rng.random((n, din))
    t = rng.normal(size=(n, dout)) * 0.1
    mask = (rng.random((n, dout)) < 0.8).astype(float)
    l64, g64 = mlp_ref.loss_grad(p, x, t, mask, din, dh, dout)
    l16, g16 = mlp_ref.loss_grad_bf16(p, x, t, mask, din, dh, dout)
    assert l16 == pytest.approx(l64, rel=2e-2)
    assert np.linalg.norm(g16 - g64) < 2e-2 * np.linalg.norm(g64)
    # every weight matrix separately, so a wrong block cannot hide behind the others
    for a, b in ((0, dh * din), (dh * din + dh, dh * din + dh + dh * dh),
                 (npar - dout - dout * dh, npar - dout)):
        assert np.linalg.norm(g16[a:b] - g64[a:b]) < 3e-2 * np.linalg.norm(g64[a:b])
    y16, _ = mlp_ref.forward_bf16(p, x, din, dh, dout)
    y64, _ = mlp_ref.forward(p, x, din, dh, dout)
    assert np.abs(y16 - y64).max() < 2e-2 * np.abs(y64).max()


def test_mean_field_block_closed_form_and_invariants():
    from oracle import mlp_ref

    # q = 2: one state per locus, C = [[c11, c12], [c12, c22]], J = c12 / (c11 c22 - c12^2)
    N = np.array([[30, 10], [5, 55]])
    B, lam = 100, 0.5
    f = N / B
    fij = (1 - lam) * f + lam / 4
    fr = (1 - lam) * f.sum(1) + lam / 2
    fj = (1 - lam) * f.sum(0) + lam / 2
    c11, c22 = fr[0] * (1 - fr[0]), fj[0] * (1 - fj[0])
    c12 = fij[0, 0] - fr[0] * fj[0]
    J = mlp_ref.mean_field_block(N, B, 2)
    assert J[0, 0] == pytest.approx(c12 / (c11 * c22 - c12 ** 2), rel=1e-12)
    assert J[1, :].tolist() == [0.0, 0.0] and J[:, 1].tolist() == [0.0, 0.0]
    # independent pair (joint = product of marginals) without pseudocount: no coupling
    # (the pseudocount mixes in a uniform table, which is not a product of these marginals)
    rng = np.random.default_rng(3)
    for q in (3, 4, 5):
        a = rng.integers(1, 9, q)
        b = rng.integers(1, 9, q)
        Nind = np.outer(a, b)
        assert np.abs(mlp_ref.mean_field_block(Nind, int(Nind.sum()), q, pseudocount=0.0)).max() < 1e-9
        assert np.abs(mlp_ref.mean_field_block(Nind, int(Nind.sum()), q)).max() > 1e-3
        # swapping the pair transposes the block
        Nr = rng.integers(0, 40, (q, q))
        Jrj = mlp_ref.mean_field_block(Nr, int(Nr.sum()), q)
        Jjr = mlp_ref.mean_field_block(Nr.T, int(Nr.sum()), q)
        np.testing.assert_allclose(Jjr, Jrj.T, rtol=1e-10, atol=1e-12)
    # a positively associated pair gets a positive coupling on the shared state
    Npos = np.array([[40, 2, 1], [2, 40, 1], [1, 1, 12]])
    assert mlp_ref.mean_field_block(Npos, int(Npos.sum()), 3)[0, 0] > 0


# ---------------------------------------------------------- norm of mean (float path)

def test_float_gauge_hand_computed_block():
    # quarters are exact in float: row means (0.75, 0.25, 0, 1), column means
    # (0.25, 0.75, 0, 1), their mean 0.5; G[i][k] = M[i][k] - rowmean[k] - colmean[i] + 0.5
    M = np.array([[1, 2, 0, 0], [0, 1, 0, 0], [0, 0, 0, 0], [0, 0, 0, 4]], dtype=np.float32)
    want = np.array([[0.5, 2.0, 0.25, -0.75],
                     [-1.0, 0.5, -0.25, -1.25],
                     [-0.25, 0.25, 0.5, -0.5],
                     [-1.25, -0.75, -0.5, 2.5]], dtype=np.float32)
    np.testing.assert_array_equal(orc.ising_gauge_float(M), want)
    # gap (state 3) excluded: (0.25 + 4 + 0.0625) + (1 + 0.25 + 0.0625) + (0.0625 + 0.0625 + 0.25) = 6
    out = orc.nom_pair(M, M)
    assert out[1] == np.float32(np.sqrt(np.float32(6.0)))


@pytest.mark.parametrize("q", [3, 4, 5])
def test_float_gauge_agrees_with_the_fp64_gauge(q):
    rng = np.random.default_rng(q)
    for _ in range(50):
        M = rng.normal(scale=0.1, size=(q, q)).astype(np.float32)
        g32 = orc.ising_gauge_float(M).astype(np.float64)
        g64 = orc.ising_gauge(M.astype(np.float64))
        assert np.abs(g32 - g64).max() <= 8 * np.finfo(np.float32).eps * max(1.0, np.abs(M).max())


def test_norm_of_mean_invariants():
    rng = np.random.default_rng(2)
    for _ in range(100):
        A = rng.normal(scale=0.1, size=(4, 4)).astype(np.float32)
        B = rng.normal(scale=0.1, size=(4, 4)).astype(np.float32)
        same = orc.nom_pair(A, A)
        assert same[0] == same[1] == same[2]  # the mean of a block with itself is the block
        jm, jij, jji = (float(v) for v in orc.nom_pair(A, B))
        assert jm <= 0.5 * (jij + jji) * (1 + 1e-6)  # the norm of a mean is at most the mean norm
        assert (jij, jji) == tuple(float(v) for v in orc.nom_pair(B, A)[[2, 1]])


def test_nom_block_orientation():
    # J_r[n](a=1, s=2) = 7 (a: state at n, s: state at r) lands at [state at the larger
    # locus][state at the smaller] of the stored block, for n above and below r
    L, q = 20, 4
    for r, n in ((3, 17), (17, 3)):
        x = np.zeros(orc.nparams(L, q))
        x[orc.ref_index(L, q, n, 1, 2)] = 7.0
        M = orc.nom_block(x, L, q, r, n)
        big, small = (1, 2) if n > r else (2, 1)  # (state at the larger, at the smaller)
        assert M[big, small] == 7.0 and np.count_nonzero(M) == 1
