"""Edge cases of the GPU path, through the C ABI: an empty locus list, repeated loci, the
smallest alignments (one sequence; two loci, where each locus' objective is a
multinomial logistic regression on the other) and identical sequences."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as orc
import superdca_amd as sd
from superdca_amd import synth

pytestmark = pytest.mark.gpu


def test_empty_locus_list(c1):
    st = c1["states"]
    with sd.Context(st) as ctx:
        w, _ = ctx.reweight(0.9)
        ctx.set_weights(w)
        rows, stats, _, rc = ctx.solve(np.zeros(0, dtype=np.int32))
    assert rc == sd.SDCA_OK
    assert rows.shape == (0, st.shape[1]) and stats["fval"].size == 0


def test_repeated_loci_give_identical_results(c1):
    # the same locus in four slots of one batch: every kernel treats a slot on its own and
    # reduces in a fixed order, so the copies agree bit for bit
    st = c1["states"]
    loci = np.array([5, 5, 17, 5], dtype=np.int32)
    with sd.Context(st, batch_loci=4, precision=sd.PREC_SPLIT_BF16) as ctx:
        w, _ = ctx.reweight(0.9)
        ctx.set_weights(w)
        rows, stats, _, rc = ctx.solve(loci)
    assert rc == sd.SDCA_OK
    for k in (1, 3):
        np.testing.assert_array_equal(rows[k], rows[0])
        assert stats["fval"][k] == stats["fval"][0] and stats["nfeval"][k] == stats["nfeval"][0]


def test_one_sequence_and_identical_sequences_reweight():
    with sd.Context(np.array([[0, 1, 2, 0]], dtype=np.uint8)) as ctx:
        w, iw = ctx.reweight(0.9)
    assert w.tolist() == [1.0] and iw.tolist() == [1]
    st = np.array([[0, 1, 2, 0], [0, 1, 2, 0], [1, 0, 2, 1]], dtype=np.uint8)
    with sd.Context(st) as ctx:
        w, iw = ctx.reweight(0.9)
    assert w.tolist() == [0.5, 0.5, 1.0] and iw.tolist() == [2, 2, 1]


def test_two_locus_alignment_against_oracle():
    d = synth.generate(B=300, N=2, clades=4, pairs=1, seed=77)
    st, _ = sd.prepare_alignment(d["chars"])
    assert st.shape[1] == 2
    w, _ = orc.weights(st, 0.9)
    _, lh, lJ = orc.lambdas(w)
    loci = np.array([0, 1], dtype=np.int32)
    ref, _, so = orc.solve_loci(st, w, lh, lJ, loci, gtol=1e-9, max_iter=10000)
    with sd.Context(st, grad_threshold=1e-6, batch_loci=2) as ctx:
        ctx.set_weights(w)
        rows, stats, _, rc = ctx.solve(loci)
    assert rc in (sd.SDCA_OK, sd.SDCA_E_NOCONV)
    assert np.all(stats["gnorm"] < 1e-4)
    np.testing.assert_allclose(stats["fval"], so["fval"], rtol=1e-6)
    # S(0, 1) and S(1, 0): the coupling score of the only pair, from each side
    np.testing.assert_allclose([rows[0, 1], rows[1, 0]], [ref[0, 1], ref[1, 0]], rtol=1e-3)
    assert rows[0, 0] == 0.0 and rows[1, 1] == 0.0
