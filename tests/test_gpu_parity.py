"""GPU parity: the HIP path (through the C ABI of libsdca.so) against the FP64 oracle.

Bars: reweighting bit-exact; objective value and gradient within the stated FP32 /
split-BF16 / fine-level tolerances; solved scores against the oracle's optimum within
1e-4 relative.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as orc
import superdca_amd as sd
from superdca_amd import synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def c2():
    d = synth.workload("C2")
    states, loci = sd.prepare_alignment(d["chars"])
    return {"states": states, "pairs": d["pairs"]}


# ------------------------------------------------------------ reweighting

@pytest.mark.parametrize("theta", [0.9, 0.95, 0.5, 1.0])
def test_reweight_bit_exact_c1(c1, theta):
    st = c1["states"]
    with sd.Context(st) as ctx:
        w, iw = ctx.reweight(theta)
    wo, iwo = orc.weights(st, theta, nthreads=8)
    np.testing.assert_array_equal(iw, iwo)
    np.testing.assert_array_equal(w, wo)


def test_reweight_bit_exact_with_multiplicity(c1):
    st = c1["states"][:, :700]
    mult = np.random.default_rng(4).integers(1, 5, st.shape[0]).astype(np.uint32)
    with sd.Context(st) as ctx:
        w, iw = ctx.reweight(0.9, mult)
    wo, iwo = orc.weights(st, 0.9, mult=mult, nthreads=8)
    np.testing.assert_array_equal(iw, iwo)
    np.testing.assert_array_equal(w, wo)


def test_reweight_bit_exact_c2(c2):
    st = c2["states"]
    with sd.Context(st) as ctx:
        w, iw = ctx.reweight(0.9)
    wo, iwo = orc.weights(st, 0.9, nthreads=16)
    np.testing.assert_array_equal(iw, iwo)
    np.testing.assert_array_equal(w, wo)


# ---------------------------------------------------- objective and gradient

EVAL_LOCI = [0, 1, 17, 500, 991, 999]  # 1000 % 16 = 8: 992..999 form the short last block


# FP32, split BF16 (hi + lo) and the fine level (three BF16 parts, accurate exp and log:
# the evaluation of a context with a coupling threshold)
@pytest.mark.parametrize("precision,jthr,ftol,gtol", [(sd.PREC_FP32, 0.0, 2e-6, 2e-5),
                                                      (sd.PREC_SPLIT_BF16, 0.0, 2e-5, 2e-4),
                                                      (sd.PREC_SPLIT_BF16, 1e-6, 2e-6, 2e-5)])
def test_eval_batch_matches_oracle(c1, precision, jthr, ftol, gtol):
    st = c1["states"]
    B, L = st.shape
    w, _ = orc.weights(st, 0.9, nthreads=8)
    rng = np.random.default_rng(12)
    n = orc.nparams(L, 4)
    x = rng.normal(size=(len(EVAL_LOCI), n)) * 0.05
    with sd.Context(st, precision=precision, batch_loci=4, j_threshold=jthr) as ctx:
        beff, lh, lJ = ctx.set_weights(w)
        f, g = ctx.eval_batch(EVAL_LOCI, x)
    for k, r in enumerate(EVAL_LOCI):
        fo, go = orc.objective(st, r, w, lh, lJ, x[k])
        assert f[k] == pytest.approx(fo, rel=ftol), r
        err = np.abs(g[k] - go).max() / np.abs(go).max()
        assert err < gtol, (r, err)


def test_eval_batch_cold_point_is_exact_softmax(c1):
    # at x = 0 every E_b is 0: f = B_eff * log(q) exactly up to FP32 weight rounding
    st = c1["states"]
    w, _ = orc.weights(st, 0.9, nthreads=8)
    L = st.shape[1]
    with sd.Context(st) as ctx:
        beff, lh, lJ = ctx.set_weights(w)
        f, g = ctx.eval_batch([3], np.zeros((1, orc.nparams(L, 4))))
    assert f[0] == pytest.approx(beff * np.log(4.0), rel=1e-6)
    fo, go = orc.objective(st, 3, w, lh, lJ, np.zeros(orc.nparams(L, 4)))
    np.testing.assert_allclose(g[0], go, atol=1e-5)


# ------------------------------------------------------------------ solve

SOLVE_LOCI = np.arange(0, 1000, 25)  # 40 loci spread over the alignment


@pytest.fixture(scope="module")
def c1_optimum(c1):
    st = c1["states"]
    w, _ = orc.weights(st, 0.9, nthreads=8)
    beff, lh, lJ = orc.lambdas(w)
    rows, X, stats = orc.solve_loci(st, w, lh, lJ, SOLVE_LOCI, gtol=1e-9, max_iter=10000, nthreads=8, want_x=True)
    assert np.all((stats["status"] == 0) | ((stats["status"] == 2) & (stats["gnorm"] < 1e-7)))
    return {"w": w, "rows": rows, "x": X, "stats": stats, "lam": (lh, lJ)}


def score_errors(rows, ref):
    mask = np.ones_like(ref, dtype=bool)
    for k, r in enumerate(SOLVE_LOCI):
        mask[k, r] = False
    big = mask & (ref > np.quantile(ref[mask], 0.9))
    rel = np.abs(rows - ref)[big] / ref[big]
    return np.abs(rows - ref)[mask].max(), rel.max()


def test_gradient_at_the_optimum_is_below_the_coupling_threshold(c1, c1_optimum):
    # at the oracle's optimum (||g||_inf < 1e-7) the GPU gradient is its own error: it must
    # lie well below the coupling threshold 3e-6 the accurate solves stop at
    st = c1["states"]
    L, q = st.shape[1], 4
    hidx = L * q * q + np.arange(q)
    for precision, jthr in ((sd.PREC_SPLIT_BF16, 1e-6), (sd.PREC_SPLIT_BF16, 0.0), (sd.PREC_FP32, 1e-6)):
        with sd.Context(st, precision=precision, batch_loci=8, j_threshold=jthr) as ctx:
            ctx.set_weights(c1_optimum["w"])
            f, g = ctx.eval_batch(SOLVE_LOCI[:8], c1_optimum["x"][:8])
        jm = np.ones(g.shape[1], dtype=bool)
        jm[hidx] = False
        gJ, gh = np.abs(g[:, jm]).max(), np.abs(g[:, hidx]).max()
        frel = np.abs(f - c1_optimum["stats"]["fval"][:8]) / np.abs(c1_optimum["stats"]["fval"][:8])
        print(f"precision {precision} j {jthr:g}: |g_J| {gJ:.2e}, |g_h| {gh:.2e} at the optimum, f rel {frel.max():.2e}")
        assert gJ < 1e-6 and gh < 1e-6
        assert frel.max() < 1e-6


# accurate solves: SuperDCA's rule (||g||_inf < 1e-3) plus the coupling threshold
# ||g_J||_inf < 3e-6, in FP32, at split precision alone and with the fine level; a locus may
# end on a line-search stall below the threshold's neighbourhood
ACCURATE = [pytest.param(dict(formulation=sd.FORM_PRIMAL, precision=sd.PREC_FP32), id="fp32"),
            pytest.param(dict(formulation=sd.FORM_PRIMAL, precision=sd.PREC_SPLIT_BF16, fine_factor=0), id="split"),
            pytest.param(dict(formulation=sd.FORM_PRIMAL, precision=sd.PREC_SPLIT_BF16, fine_factor=10),
                         id="split+fine"),
            pytest.param(dict(formulation=sd.FORM_GRAM, j_threshold=1e-5), id="gram")]


@pytest.mark.parametrize("kw", ACCURATE)
def test_solve_accurate_matches_oracle_optimum(c1, c1_optimum, kw):
    st = c1["states"]
    with sd.Context(st, batch_loci=64, **{"j_threshold": 3e-6, **kw}) as ctx:
        ctx.set_weights(c1_optimum["w"])
        rows, stats, X, rc = ctx.solve(SOLVE_LOCI, want_x=True)
    assert rc in (sd.SDCA_OK, sd.SDCA_E_NOCONV)
    ok = (stats["status"] == sd.ST_CONVERGED) | ((stats["status"] == sd.ST_LS_FAIL) & (stats["gnorm_J"] < 3e-5))
    assert ok.all(), stats
    np.testing.assert_allclose(stats["fval"], c1_optimum["stats"]["fval"], rtol=1e-6)
    for k, r in enumerate(SOLVE_LOCI):
        assert rows[k, r] == 0.0
    abs_err, rel_top = score_errors(rows, c1_optimum["rows"])
    xerr = np.abs(X - c1_optimum["x"]).max() / np.abs(c1_optimum["x"]).max()
    print(f"{kw}: {int((stats['status'] != 0).sum())} line-search stops, scores max abs err {abs_err:.3e}, "
          f"max rel err (top decile) {rel_top:.3e}, parameters max err {xerr:.3e} of max |x|")
    assert rel_top < 1e-4
    if kw.get("formulation") == sd.FORM_GRAM:
        # FP64 evaluations: f at the optimum to its rounding.  The parameters are as close as
        # the stopping point allows: the objective is 2 min(lambda_h, lambda_J)-strongly
        # convex, so ||x - x*||_2 <= ||g(x)||_2 / (2 min(lambda_h, lambda_J)) with the reported
        # 2-norm of the final gradient; the couplings are formed once with FP32 products
        # (1e-6 of ||x*||_2 on top).  A wrong layout, sign or scale misses this by orders.
        np.testing.assert_allclose(stats["fval"], c1_optimum["stats"]["fval"], rtol=1e-10)
        mu = 2.0 * min(c1_optimum["lam"])
        dist = np.linalg.norm(X - c1_optimum["x"], axis=1)
        bound = stats["gnorm"] / mu + 1e-6 * np.linalg.norm(c1_optimum["x"], axis=1)
        print(f"  parameters: largest distance / bound {np.max(dist / bound):.3f}")
        assert np.all(dist <= bound)


def test_reference_mode_threshold_and_refill(c1, c1_optimum):
    # default gradient threshold 1e-3 (plmDCA_options.cpp:61) with fewer slots than loci:
    # converged slots are refilled from the queue
    st = c1["states"]
    with sd.Context(st, batch_loci=7) as ctx:
        ctx.set_weights(c1_optimum["w"])
        rows, stats, _, rc = ctx.solve(SOLVE_LOCI)
    assert rc == sd.SDCA_OK
    assert np.all(stats["gnorm"] < 1e-3)
    np.testing.assert_allclose(stats["fval"], c1_optimum["stats"]["fval"], rtol=1e-4)


def test_warm_start_reaches_cold_optimum(c1, c1_optimum):
    st = c1["states"]
    rng = np.random.default_rng(5)
    # a start near the optimum (10 % multiplicative perturbation of every parameter)
    warm = c1_optimum["x"] * (1.0 + 0.1 * rng.normal(size=c1_optimum["x"].shape))
    with sd.Context(st, grad_threshold=1e-5, batch_loci=64, formulation=sd.FORM_PRIMAL) as ctx:
        ctx.set_weights(c1_optimum["w"])
        _, sw, _, rc = ctx.solve(SOLVE_LOCI, warm=warm)
        _, sc, _, rc2 = ctx.solve(SOLVE_LOCI)
    assert rc == sd.SDCA_OK and rc2 == sd.SDCA_OK
    np.testing.assert_allclose(sw["fval"], sc["fval"], rtol=1e-6)
    assert np.median(sw["iter"]) < np.median(sc["iter"])


def test_gram_graph_rounds_match_launches(c1, c1_optimum):
    # each slot group's round replayed from a captured HIP graph (fixed grids over all of the
    # group's slots, idle ones skipped inside the kernels) against rounds launched over the
    # active slots: 80 loci through 64 slots (two groups, refills in both, idle slots at
    # the end), then 80 more through the same workspace, so the graphs are replayed across
    # solves with the host's own work and synchronisations between the replays; every
    # result must be identical
    st = c1["states"]
    first, second = np.arange(1, 1000, 12)[:80], np.arange(5, 1000, 12)[:80]
    out = {}
    for graph in (0, 1):
        with sd.Context(st, formulation=sd.FORM_GRAM, grad_threshold=1e-4, j_threshold=1e-5, batch_loci=64,
                        graph_rounds=graph) as ctx:
            ctx.set_weights(c1_optimum["w"])
            r1, s1, _, rc1 = ctx.solve(first)
            r2, s2, _, rc2 = ctx.solve(second)
            tm = ctx.timing()
        assert rc1 == sd.SDCA_OK and rc2 == sd.SDCA_OK, graph
        out[graph] = (r1, s1, r2, s2, tm)
    for k in (0, 2):
        np.testing.assert_array_equal(out[1][k], out[0][k])
    for k in (1, 3):
        for f in ("fval", "gnorm", "nfeval", "iter", "status"):
            np.testing.assert_array_equal(out[1][k][f], out[0][k][f])
    assert out[1][4]["n_evals"] == out[0][4]["n_evals"]
    assert out[1][4]["gemm_grad_ms"] == 0.0 and out[1][4]["lbfgs_ms"] > 0.0  # a replayed round is timed whole


def test_cli_end_to_end(tmp_path):
    d = synth.generate(B=150, N=120, clades=6, pairs=4, seed=77)
    fa = tmp_path / "toy.fasta"
    synth.write_fasta(str(fa), d["chars"])
    exe = os.path.join(ROOT, "superdca_amd", "sdca")
    # chunks of 37 loci: the scores reach the file-backed store from several solve calls
    p = subprocess.run([exe, "-v", "--chunk-loci", "37", str(fa)], capture_output=True, text=True, cwd=tmp_path,
                       timeout=300)
    assert p.returncode == 0, p.stderr
    out = tmp_path / "toy.SuperDCA_couplings.1-based.all"
    lines = out.read_text().splitlines()
    states, loci = sd.prepare_alignment(d["chars"])
    L = states.shape[1]
    assert len(lines) == L * (L - 1) // 2
    first = [int(t.split()[1]) > int(t.split()[2]) for t in lines]
    assert all(first)
    # the numbers of the library called directly (to the last digits an autotuned split-K moves)
    with sd.Context(states) as ctx:
        w, _ = ctx.reweight(0.9)
        ctx.set_weights(w)
        rows, _, _, _ = ctx.solve(np.arange(L))
    ref = tmp_path / "direct.all"
    sd.write_couplings(str(ref), rows, loci, 1)
    assert_same_couplings(lines, ref.read_text().splitlines())


# ------------------------------------------------------------ GEMM configurations

@pytest.mark.parametrize("cfg", [
    dict(autotune=1),
    dict(gemm_ksplit=1, gemm_gm_energy=1, gemm_gm_grad=1, gemm_tn_energy=128, gemm_tn_grad=128),
    dict(gemm_ksplit=3, gemm_gm_energy=2, gemm_gm_grad=4, gemm_tn_energy=256, gemm_tn_grad=256),
    dict(gemm_ksplit=7, gemm_gm_energy=64, gemm_gm_grad=3, gemm_tn_energy=256, gemm_tn_grad=128),
    dict(gemm_ksplit=2, gemm_gm_energy=1, gemm_gm_grad=8, gemm_tn_energy=128, gemm_tn_grad=256)])
@pytest.mark.parametrize("precision,jthr,ftol,gtol", [(sd.PREC_FP32, 0.0, 2e-6, 2e-5),
                                                      (sd.PREC_SPLIT_BF16, 0.0, 2e-5, 2e-4),
                                                      (sd.PREC_SPLIT_BF16, 1e-6, 2e-6, 2e-5)])
def test_gemm_configurations_match_oracle(c1, cfg, precision, jthr, ftol, gtol):
    # every tile walk / split-K the autotuner may choose computes the same objective, with
    # two MFMA passes (split) or three (the fine level of a context with a coupling
    # threshold); 20 slots x q = 80 rows pad to one 128-row tile, 48 slots to two
    st = c1["states"]
    w, _ = orc.weights(st, 0.9, nthreads=8)
    n = orc.nparams(st.shape[1], 4)
    loci = np.arange(3, 1000, 21)[:48]
    x = np.random.default_rng(31).normal(size=(loci.size, n)) * 0.05
    with sd.Context(st, precision=precision, batch_loci=48, j_threshold=jthr, **cfg) as ctx:
        beff, lh, lJ = ctx.set_weights(w)
        f, g = ctx.eval_batch(loci, x)
        got = ctx.gemm_config()
        kx = sd.lib().sdca_num_columns(ctx._h)
    if "gemm_ksplit" in cfg:
        assert not got["tuned"]
        assert got["ksplit"] == cfg["gemm_ksplit"]
        assert (got["gm_energy"], got["gm_grad"]) == (cfg["gemm_gm_energy"], cfg["gemm_gm_grad"])
        assert (got["tn_energy"], got["tn_grad"]) == (cfg["gemm_tn_energy"], cfg["gemm_tn_grad"])
    else:
        assert got["tuned"]
    assert got["kchunk"] > 0 and got["ksplit"] * got["kchunk"] > kx  # K_x columns + the field column
    for k in (0, 17, 47):
        fo, go = orc.objective(st, int(loci[k]), w, lh, lJ, x[k])
        assert f[k] == pytest.approx(fo, rel=ftol), (k, got)
        err = np.abs(g[k] - go).max() / np.abs(go).max()
        assert err < gtol, (k, err, got)


@pytest.mark.parametrize("precision", [sd.PREC_FP32, sd.PREC_SPLIT_BF16])
def test_pipelined_slot_groups_match_single_group(c1, precision):
    # 80 loci through 64 slots: two groups of 32 slots (128 GEMM rows each) alternate on two
    # streams, with refills in both; every locus must reach the same optimum as with one group
    # (and the two-group layout must stay usable although one group is the default)
    st = c1["states"]
    w, _ = orc.weights(st, 0.9, nthreads=8)
    loci = np.arange(0, 1000, 12)[:80]
    out = {}
    for groups in (1, 2):
        with sd.Context(st, precision=precision, grad_threshold=1e-4, batch_loci=64, pipeline_groups=groups,
                        formulation=sd.FORM_PRIMAL) as ctx:
            ctx.set_weights(w)
            rows, stats, _, rc = ctx.solve(loci)
        assert rc == sd.SDCA_OK, groups
        assert np.all(stats["gnorm"] < 1e-4)
        out[groups] = (rows, stats)
    np.testing.assert_allclose(out[2][1]["fval"], out[1][1]["fval"], rtol=1e-6)
    big = out[1][0] > np.quantile(out[1][0], 0.9)
    rel = np.abs(out[2][0] - out[1][0])[big] / out[1][0][big]
    assert rel.max() < (2e-3 if precision == sd.PREC_SPLIT_BF16 else 1e-3), rel.max()


def test_coarse_phase_reaches_the_full_precision_optimum(c1):
    # split-BF16 with the coarse phase (BF16 high parts only while ||g||_inf >= factor *
    # threshold) off, at the default factor, and so large that every start point is
    # already below it (re-evaluated in full at once); 80 loci through 64 slots, so coarse
    # and full-precision slots share rounds and refills
    st = c1["states"]
    w, _ = orc.weights(st, 0.9, nthreads=8)
    loci = np.arange(3, 1000, 12)[:80]
    out = {}
    for factor in (0, 4, 1000000):
        with sd.Context(st, precision=sd.PREC_SPLIT_BF16, grad_threshold=1e-4, batch_loci=64,
                        coarse_factor=factor, formulation=sd.FORM_PRIMAL) as ctx:
            ctx.set_weights(w)
            rows, stats, _, rc = ctx.solve(loci)
        assert rc == sd.SDCA_OK, factor
        assert np.all(stats["status"] == sd.ST_CONVERGED), factor
        assert np.all(stats["gnorm"] < 1e-4), factor
        out[factor] = (rows, stats)
    for factor in (4, 1000000):
        np.testing.assert_allclose(out[factor][1]["fval"], out[0][1]["fval"], rtol=1e-6)
        big = out[0][0] > np.quantile(out[0][0], 0.9)
        rel = np.abs(out[factor][0] - out[0][0])[big] / out[0][0][big]
        assert rel.max() < 2e-3, (factor, rel.max())
    print("coarse phase evaluations", {f: int(o[1]["nfeval"].sum()) for f, o in out.items()})


@pytest.mark.parametrize("precision,ftol,gtol", [(sd.PREC_FP32, 2e-6, 2e-5), (sd.PREC_SPLIT_BF16, 2e-5, 2e-4)])
def test_eval_and_solve_q5_envelope(c1, precision, ftol, gtol):
    # the genome-scale envelope {A,C,G,T,-}: q_env = 5, gap = state 4, with an envelope
    # state that no locus of C1 observes (three alleles at most)
    st5, _ = sd.prepare_alignment(c1["chars"], q=5)
    assert st5.max() <= 4
    w, _ = orc.weights(st5, 0.9, nthreads=8)
    n = orc.nparams(st5.shape[1], 5)
    loci = [0, 17, 500, 999]
    x = np.random.default_rng(41).normal(size=(len(loci), n)) * 0.05
    with sd.Context(st5, q=5, precision=precision, batch_loci=4, formulation=sd.FORM_PRIMAL) as ctx:
        beff, lh, lJ = ctx.set_weights(w)
        f, g = ctx.eval_batch(loci, x)
        rows, stats, _, rc = ctx.solve(loci)
    for k, r in enumerate(loci):
        fo, go = orc.objective(st5, r, w, lh, lJ, x[k], q=5)
        assert f[k] == pytest.approx(fo, rel=ftol), r
        assert np.abs(g[k] - go).max() / np.abs(go).max() < gtol, r
    assert rc == sd.SDCA_OK and np.all(stats["gnorm"] < 1e-3)
    ro, _, so = orc.solve_loci(st5, w, lh, lJ, np.array(loci), q=5, gtol=1e-7, max_iter=10000, nthreads=4)
    np.testing.assert_allclose(stats["fval"], so["fval"], rtol=1e-5)


def assert_same_couplings(got: list, ref: list, atol: float = 2e-7):
    """Two coupling files with the same pairs in the same order and scores within atol: each
    process autotunes its own split-K, which changes the FP32 summation order of the
    energies and so the last digits of a solve stopped at SuperDCA's rule."""
    assert len(got) == len(ref)
    a = [t.split() for t in got]
    b = [t.split() for t in ref]
    assert [t[1:] for t in a] == [t[1:] for t in b]
    va = np.array([float(t[0]) for t in a])
    vb = np.array([float(t[0]) for t in b])
    assert np.abs(va - vb).max() <= atol, np.abs(va - vb).max()
