"""The subspace warm start of the Gram formulation (sdca_opts.subspace_rank) on the GPU:
the same optimum as the cold start, fewer full evaluations, deterministic, and skipped
where it does not apply.  Its accuracy against the FP64 oracle is covered by the
"gram+sub" cases of tests/test_gpu_scores.py (C2 whole alignment, C4 rows)."""
import numpy as np
import pytest

import superdca_amd as sd
from superdca_amd import synth

pytestmark = pytest.mark.gpu

GRAM = dict(formulation=sd.FORM_GRAM, grad_threshold=1e-3, j_threshold=1e-5)


@pytest.fixture(scope="module")
def c2_q5():
    st, _ = sd.prepare_alignment(synth.workload("C2")["chars"], q=5)
    with sd.Context(st, q=5) as ctx:
        w, _ = ctx.reweight(0.9)
    return st, w


def solve(st, w, loci, **kw):
    with sd.Context(st, q=5, batch_loci=128, **{**GRAM, **kw}) as ctx:
        ctx.set_weights(w)
        rows, stats, x, rc = ctx.solve(loci, want_x=True)
        tm = ctx.timing()
    assert rc == sd.SDCA_OK, rc
    assert np.all(stats["status"] == sd.ST_CONVERGED)
    return rows, stats, x, tm


def test_subspace_same_optimum_fewer_full_evaluations(c2_q5):
    st, w = c2_q5
    loci = np.random.default_rng(3).choice(st.shape[1], 300, replace=False).astype(np.int32)
    rc_, sc, xc, _ = solve(st, w, loci, subspace_rank=0)
    rs, ss, xs, tm = solve(st, w, loci, subspace_rank=128)
    # strictly convex: both starts reach the optimum (FP64 evaluations)
    np.testing.assert_allclose(ss["fval"], sc["fval"], rtol=1e-10)
    top = np.argsort(-rc_, axis=1)[:, :50]
    a, b = np.take_along_axis(rc_, top, 1), np.take_along_axis(rs, top, 1)
    assert np.max(np.abs(a - b) / a) < 1e-4
    assert np.abs(xs - xc).max() < 1e-4 * np.abs(xc).max()
    assert np.all(ss["nfeval_sub"] > 0) and np.all(sc["nfeval_sub"] == 0)
    assert tm["n_sub_evals"] > 0 and tm["subspace_ms"] > 0
    print(f"C2 300 loci: full evaluations cold {sc['nfeval'].sum()} warm {ss['nfeval'].sum()} "
          f"(+ {ss['nfeval_sub'].sum()} restricted), median {np.median(sc['nfeval']):.0f} / "
          f"{np.median(ss['nfeval']):.0f}")
    assert ss["nfeval"].sum() < 0.7 * sc["nfeval"].sum()


def test_subspace_solve_is_deterministic(c2_q5):
    st, w = c2_q5
    loci = np.arange(0, st.shape[1], 37, dtype=np.int32)
    r1, s1, x1, _ = solve(st, w, loci, subspace_rank=128)
    r2, s2, x2, _ = solve(st, w, loci, subspace_rank=128)
    np.testing.assert_array_equal(r1, r2)
    np.testing.assert_array_equal(x1, x2)
    for k in ("fval", "nfeval", "nfeval_sub", "iter"):
        np.testing.assert_array_equal(s1[k], s2[k])


def test_subspace_results_independent_of_locus_order(c2_q5):
    """The full solves start in the order of their restricted evaluations, and both phases
    assign loci to slots as slots free up: a permuted locus list gives every locus the
    same result, bit for bit."""
    st, w = c2_q5
    loci = np.arange(0, st.shape[1], 29, dtype=np.int32)
    perm = np.random.default_rng(5).permutation(loci.size)
    r1, s1, x1, _ = solve(st, w, loci, subspace_rank=128)
    r2, s2, x2, _ = solve(st, w, loci[perm], subspace_rank=128)
    np.testing.assert_array_equal(r1[perm], r2)
    np.testing.assert_array_equal(x1[perm], x2)
    for k in ("fval", "nfeval", "nfeval_sub", "iter"):
        np.testing.assert_array_equal(s1[k][perm], s2[k])


@pytest.mark.parametrize("groups", [1, 4])
def test_slot_groups_give_identical_results(c2_q5, groups):
    """One, two (the default) or four slot groups on as many streams: every locus' rounds
    depend on its own slot only, so the results are bit-identical.  With 295 loci in 256
    slots, two groups run the restricted rounds in four (64 slots each), one group in one."""
    st, w = c2_q5
    loci = np.arange(0, st.shape[1], 17, dtype=np.int32)
    out = []
    for g in (2, groups):
        with sd.Context(st, q=5, batch_loci=256, pipeline_groups=g, subspace_rank=128, **GRAM) as ctx:
            ctx.set_weights(w)
            rows, stats, _, rc = ctx.solve(loci)
        assert rc == sd.SDCA_OK, rc
        out.append((rows, stats))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for k in ("fval", "nfeval", "nfeval_sub", "iter"):
        np.testing.assert_array_equal(out[0][1][k], out[1][1][k])


@pytest.mark.parametrize("slots", [257, 300])
def test_restricted_groups_with_uneven_slot_counts(c2_q5, slots):
    """Slot counts that four restricted groups do not divide evenly (257: three groups of 65
    slots and one of 62; 300: four of 75): the same results as one group."""
    st, w = c2_q5
    loci = np.arange(0, st.shape[1], 13, dtype=np.int32)  # 385 loci
    out = []
    for g in (2, 1):
        with sd.Context(st, q=5, batch_loci=slots, pipeline_groups=g, subspace_rank=128, **GRAM) as ctx:
            ctx.set_weights(w)
            rows, stats, _, rc = ctx.solve(loci)
        assert rc == sd.SDCA_OK, rc
        assert stats["nfeval_sub"].sum() > 0
        out.append((rows, stats))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for k in ("fval", "nfeval", "nfeval_sub", "iter"):
        np.testing.assert_array_equal(out[0][1][k], out[1][1][k])


def test_subspace_skipped_below_256_sequences(c1):
    """B = 200 < 2 x 128: the alignment solves from zero, bit-identical to the cold start."""
    st = c1["states"]
    loci = np.arange(0, st.shape[1], 25, dtype=np.int32)
    out = []
    for rank in (0, 128):
        with sd.Context(st, batch_loci=64, subspace_rank=rank, **GRAM) as ctx:
            w, _ = ctx.reweight(0.9)
            ctx.set_weights(w)
            rows, stats, _, rc = ctx.solve(loci)
        out.append((rows, stats))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1]["fval"], out[1][1]["fval"])
    assert np.all(out[1][1]["nfeval_sub"] == 0)


def test_subspace_skipped_when_its_basis_is_singular():
    """B = 300 sequences but only 60 loci: the identical-locus counts have too few
    directions for a 128-dimensional basis (Q^T K Q singular), so the loci start from zero,
    bit-identical to the cold start, on this call and the next."""
    d = synth.generate(B=300, N=60, clades=8, pairs=3, seed=12)
    st, _ = sd.prepare_alignment(d["chars"])
    loci = np.arange(0, st.shape[1], 7, dtype=np.int32)
    out = []
    for rank in (0, 128):
        with sd.Context(st, batch_loci=32, subspace_rank=rank, **GRAM) as ctx:
            w, _ = ctx.reweight(0.9)
            ctx.set_weights(w)
            for _ in range(2):
                rows, stats, _, rc = ctx.solve(loci)
                assert rc == sd.SDCA_OK, rc
                out.append((rows, stats))
    for rows, stats in out[1:]:
        np.testing.assert_array_equal(out[0][0], rows)
        np.testing.assert_array_equal(out[0][1]["fval"], stats["fval"])
        assert np.all(stats["nfeval_sub"] == 0)

