"""The uneven batch (superdca_amd.synth.c3_workloads: 48 alignments of different shapes):
its three smallest alignments, each with its own envelope (q_env 3, 4 and 5) and its own
context, reweighted and solved on the GPU (through the C ABI) and compared locus by locus
with the FP64 oracle's optimum.  The oracle solves in the test (about 10 s)."""
from __future__ import annotations

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
@pytest.mark.parametrize("kw", [pytest.param(dict(precision=sd.PREC_FP32), id="fp32"),
                                pytest.param(dict(precision=sd.PREC_SPLIT_BF16, fine_factor=0), id="split"),
                                pytest.param(dict(precision=sd.PREC_SPLIT_BF16, fine_factor=10), id="split+fine")])
def test_c3_loci_against_oracle_optimum(c3_cases, kw):
    for c in c3_cases:
        st, loci, ref = c["states"], c["loci"], c["rows"]
        with sd.Context(st, q=c["q"], j_threshold=3e-6, batch_loci=NLOCI, **kw) as ctx:
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
