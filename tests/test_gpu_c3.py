"""The uneven batch (superdca_amd.synth.c3_workloads: 48 alignments of different shapes).

Its three smallest alignments, each with its own envelope (q_env 3, 4 and 5) and its own
context, reweighted and solved on the GPU (through the C ABI) and compared locus by locus
with the FP64 oracle's optimum solved in the test (about 10 s); and all 48 alignments plus
a two-state gapless one solved together in one multi-alignment solve, against the oracle's
optimum committed in tests/golden/c3_fixture.npz (tests/golden/make_c3_fixture.py)."""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle as orc
import superdca_amd as sd
from superdca_amd import synth

pytestmark = pytest.mark.gpu

NLOCI = 12


@pytest.fixture(scope="module")
def c3_cases():
    cases = []
    cfgs = sorted(synth.c3_workloads(), key=lambda d: d["B"] * d["N"])[:3]
    for cfg, q in zip(cfgs, (3, 4, 5)):
        st, _ = sd.prepare_alignment(synth.generate(**cfg)["chars"], q=q)
        L = st.shape[1]
        w, iw = orc.weights(st, 0.9, nthreads=16)
        _, lh, lJ = orc.lambdas(w)
        loci = np.linspace(0, L - 1, NLOCI).astype(np.int32)
        rows, _, stats = orc.solve_loci(st, w, lh, lJ, loci, q=q, gtol=1e-9, max_iter=20000, nthreads=16)
        # a line search that fails at the FP64 rounding floor leaves the optimum to within far
        # less than the GPU thresholds below
        assert np.all((stats["status"] == 0) | ((stats["status"] == 2) & (stats["gnorm"] < 1e-6)))
        cases.append({"cfg": cfg, "q": q, "states": st, "w": w, "iw": iw, "loci": loci, "rows": rows,
                      "f": stats["fval"]})
    return cases


def test_c3_reweight_bit_exact(c3_cases):
    for c in c3_cases:
        with sd.Context(c["states"], q=c["q"]) as ctx:
            w, iw = ctx.reweight(0.9)
        np.testing.assert_array_equal(iw, c["iw"])
        np.testing.assert_array_equal(w, c["w"])


# SuperDCA's rule plus the coupling threshold ||g_J||_inf < 3e-6; bars: f 3e-6 relative
# (the FP32-summed energies of the B = 103, L = 16,890 alignment carry about 1e-6), every
# locus' top-decile scores 1e-4 relative
@pytest.mark.parametrize("kw", [
    pytest.param(dict(formulation=sd.FORM_PRIMAL, precision=sd.PREC_FP32), id="fp32"),
    pytest.param(dict(formulation=sd.FORM_PRIMAL, precision=sd.PREC_SPLIT_BF16, fine_factor=0), id="split"),
    pytest.param(dict(formulation=sd.FORM_PRIMAL, precision=sd.PREC_SPLIT_BF16, fine_factor=10), id="split+fine"),
    pytest.param(dict(formulation=sd.FORM_GRAM, j_threshold=1e-5), id="gram")])
def test_c3_loci_against_oracle_optimum(c3_cases, kw):
    for c in c3_cases:
        st, loci, ref = c["states"], c["loci"], c["rows"]
        with sd.Context(st, q=c["q"], batch_loci=NLOCI, **{"j_threshold": 3e-6, **kw}) as ctx:
            ctx.set_weights(c["w"])
            rows, stats, _, rc = ctx.solve(loci)
        assert rc in (sd.SDCA_OK, sd.SDCA_E_NOCONV), rc
        assert np.all((stats["status"] == sd.ST_CONVERGED) | ((stats["status"] == sd.ST_LS_FAIL) & (stats["gnorm"] < 1e-3)))
        frel = np.abs(stats["fval"] - c["f"]) / np.abs(c["f"])
        # per locus: the relative error of the row's top decile (locus itself excluded)
        per_locus = []
        for k, r in enumerate(loci):
            m = np.ones(ref.shape[1], dtype=bool)
            m[r] = False
            big = m & (ref[k] > np.quantile(ref[k][m], 0.9))
            per_locus.append((np.abs(rows[k] - ref[k])[big] / ref[k][big]).max())
        per_locus = np.array(per_locus)
        B, L = st.shape
        print(f"C3 seed {c['cfg']['seed']} B={B} L={L} q={c['q']} {kw}: "
              f"f rel err max {frel.max():.2e}; per-locus top-decile score rel err "
              f"{np.array2string(per_locus, precision=1, separator=' ')}")
        print(f"  status {stats['status']}, ||g_J||_inf max {stats['gnorm_J'].max():.2e}")
        assert frel.max() < 3e-6
        assert per_locus.max() < 1e-4


# ------------------------------------------------- all 48 alignments in one solve

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c3_fixture.npz")


def gapless_two_state(B=300, L=700, clades=12, seed=77):
    """A seeded two-state alignment without gaps (envelope q_env = 2, no gap state): clade
    founders of independent sites, 3 % of sites redrawn per sequence."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.05, 0.5, L)
    founders = (rng.random((clades, L)) < p).astype(np.uint8)
    st = founders[rng.integers(0, clades, B)]
    flip = rng.random((B, L)) < 0.03
    st = np.where(flip, (rng.random((B, L)) < p).astype(np.uint8), st)
    keep = (st.min(axis=0) == 0) & (st.max(axis=0) == 1)
    return np.ascontiguousarray(st[:, keep])


@pytest.mark.skipif(not os.path.exists(FIXTURE), reason="tests/golden/c3_fixture.npz not generated")
@pytest.mark.parametrize("kw", [
    pytest.param(dict(formulation=sd.FORM_PRIMAL, precision=sd.PREC_SPLIT_BF16, fine_factor=0), id="split"),
    pytest.param(dict(formulation=sd.FORM_PRIMAL, precision=sd.PREC_FP32), id="fp32"),
    pytest.param(dict(formulation=sd.FORM_GRAM, j_threshold=1e-5), id="gram")])
def test_c3_all_alignments_in_one_solve(kw):
    # the 48 alignments of the uneven batch (envelopes 3, 4 and 5) and a gapless two-state
    # alignment, their loci interleaved in one multi-alignment solve: every round's GEMM
    # launches carry one tile group per alignment in flight.  The coupling threshold is 1e-6:
    # rows of up to 20,000 entries reach down to small scores in their top 200
    fx = np.load(FIXTURE)
    cfgs = synth.c3_workloads()
    ctxs, items = [], []
    for k, cfg in enumerate(cfgs):
        q = int(fx[f"q_{k}"])
        st, _ = sd.prepare_alignment(synth.generate(**cfg)["chars"], q=q)
        assert st.shape[1] == int(fx[f"L_{k}"])
        ctx = sd.Context(st, q=q)
        w, _ = ctx.reweight(0.9)
        assert abs(w.sum() - float(fx[f"w_sum_{k}"])) <= 1e-9 * float(fx[f"w_sum_{k}"])
        ctx.set_weights(w)
        ctxs.append(ctx)
        items += [(k, int(r)) for r in fx[f"loci_{k}"]]
    st2 = gapless_two_state()
    c2 = sd.Context(st2, q=2, gap=-1)
    w2, _ = c2.reweight(0.9)
    _, lh2, lJ2 = c2.set_weights(w2)
    ctxs.append(c2)
    loci2 = np.array([0, 5, st2.shape[1] // 2, st2.shape[1] - 1], dtype=np.int32)
    items += [(len(cfgs), int(r)) for r in loci2]
    order = np.random.default_rng(48).permutation(len(items))
    alns = np.array([items[i][0] for i in order], dtype=np.int32)
    loci = np.array([items[i][1] for i in order], dtype=np.int32)
    print(f"{len(items)} loci of {len(ctxs)} alignments, q_env {sorted(set(c.q for c in ctxs))}", flush=True)
    with sd.Multi(ctxs, batch_loci=256, **{"j_threshold": 1e-6, **kw}) as mu:
        rows, stats, rc = mu.solve(alns, loci)
        tm = mu.timing()
    assert rc in (sd.SDCA_OK, sd.SDCA_E_NOCONV)
    assert np.all((stats["status"] == sd.ST_CONVERGED) | ((stats["status"] == sd.ST_LS_FAIL) & (stats["gnorm"] < 1e-3)))
    print(f"  {tm['n_evals']} rounds, {int((stats['status'] != 0).sum())} line-search stops", flush=True)
    where = {(int(a), int(r)): u for u, (a, r) in enumerate(zip(alns, loci))}
    worst_f, worst_s = 0.0, 0.0
    for k in range(len(cfgs)):
        for j, r in enumerate(fx[f"loci_{k}"]):
            u = where[(k, int(r))]
            frel = abs(stats["fval"][u] - fx[f"f_{k}"][j]) / abs(fx[f"f_{k}"][j])
            got = rows[u][fx[f"top_idx_{k}"][j]]
            rel = np.abs(got - fx[f"top_val_{k}"][j]) / fx[f"top_val_{k}"][j]
            worst_f, worst_s = max(worst_f, frel), max(worst_s, rel.max())
            if rel.max() >= worst_s:
                where_worst = (k, int(r), float(fx[f"top_val_{k}"][j].min()), float(fx[f"top_val_{k}"][j].max()))
    print(f"  worst score error at alignment, locus, row top-200 range {where_worst}")
    # the two-state alignment against the oracle in the test (no gap state: nothing excluded)
    ref, _, so = orc.solve_loci(st2, w2, lh2, lJ2, loci2, q=2, gap=-1, gtol=1e-9, max_iter=20000, nthreads=16)
    for j, r in enumerate(loci2):
        u = where[(len(cfgs), int(r))]
        frel = abs(stats["fval"][u] - so["fval"][j]) / abs(so["fval"][j])
        top = np.argsort(-ref[j])[:200]
        rel = np.abs(rows[u][top] - ref[j][top]) / ref[j][top]
        worst_f, worst_s = max(worst_f, frel), max(worst_s, rel.max())
    for c in ctxs:
        c.close()
    print(f"  worst f rel err {worst_f:.2e}, worst top-200 score rel err {worst_s:.2e}")
    assert worst_f < 3e-6
    assert worst_s < 1e-4
