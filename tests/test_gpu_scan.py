"""Two-alignment scan (SuperDCA.cpp:155-159; objective plmDCA_objective.hpp:67, output
plmDCA.hpp:672-681) through the C ABI: a locus of the first alignment fitted on the second
alignment's columns alone is the ordinary per-locus problem of the alignment made of that
locus followed by the second alignment, which the FP64 oracle solves here."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import oracle as orc
import superdca_amd as sd
from superdca_amd import synth

pytestmark = pytest.mark.gpu

L1 = 20


@pytest.fixture(scope="module")
def scan_case():
    d = synth.generate(B=200, N=90, clades=8, pairs=10, seed=505)
    st, _ = sd.prepare_alignment(d["chars"])
    st2 = np.ascontiguousarray(st[:, L1:])
    w2, _ = orc.weights(st2, 0.9)  # the second alignment's weights (plmDCA.hpp:538)
    _, lh, lJ = orc.lambdas(w2)
    return st, st2, w2, lh, lJ


def test_scan_matches_the_oracle_on_the_reduced_alignment(scan_case):
    st, st2, w2, lh, lJ = scan_case
    loci = np.array([0, 7, 19], dtype=np.int32)
    with sd.Context(st, grad_threshold=1e-6, batch_loci=3) as ctx:
        ctx.set_weights(w2)
        ctx.set_scan_split(L1)
        rows, stats, _, rc = ctx.solve(loci)
    assert rc in (sd.SDCA_OK, sd.SDCA_E_NOCONV)
    assert np.all(stats["gnorm"] < 1e-4)
    for k, r in enumerate(loci):
        red = np.ascontiguousarray(np.concatenate([st[:, r:r + 1], st2], axis=1))
        ref, _, so = orc.solve_loci(red, w2, lh, lJ, np.array([0], dtype=np.int32), gtol=1e-9, max_iter=10000)
        assert stats["fval"][k] == pytest.approx(so["fval"][0], rel=1e-7)
        assert np.all(rows[k, :L1] == 0.0)  # the first alignment is no predictor
        got, want = rows[k, L1:], ref[0, 1:]
        big = want > np.quantile(want, 0.9)
        rel = np.abs(got - want)[big] / want[big]
        print(f"scan locus {r}: f {stats['fval'][k]:.8f} vs {so['fval'][0]:.8f}, top-decile rel err {rel.max():.2e}")
        assert rel.max() < 1e-5


def test_scan_refusals(scan_case):
    st, _, w2, _, _ = scan_case
    with sd.Context(st, batch_loci=2) as ctx:
        ctx.set_weights(w2)
        with pytest.raises(sd.SdcaError):
            ctx.set_scan_split(st.shape[1])
        ctx.set_scan_split(L1)
        with pytest.raises(sd.SdcaError, match="not in the first alignment"):
            ctx.solve(np.array([L1], dtype=np.int32))
        with pytest.raises(sd.SdcaError, match="scan"):
            ctx.eval_batch([0], np.zeros((1, sd.nparams(st.shape[1], 4))))
        with pytest.raises(sd.SdcaError, match="cold start"):
            ctx.solve(np.array([0], dtype=np.int32), warm=np.zeros((1, sd.nparams(st.shape[1], 4))))
        ctx.set_scan_split(0)  # off again: the ordinary solve
        _, stats, _, rc = ctx.solve(np.array([L1], dtype=np.int32))
        assert rc == sd.SDCA_OK and stats["status"][0] == sd.ST_CONVERGED


def test_cli_scan_end_to_end(tmp_path):
    # sdca with two alignment files: the numbers of the library driven directly
    d = synth.generate(B=120, N=50, clades=6, pairs=5, seed=606)
    a, b = d["chars"][:, :15], d["chars"][:, 15:]
    for name, part in (("first.fa", a), ("second.fa", b)):
        synth.write_fasta(str(tmp_path / name), part)
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "superdca_amd", "sdca")
    p = subprocess.run([exe, "--output-weights", "first.fa", "second.fa"], capture_output=True, text=True,
                       cwd=tmp_path, timeout=300)
    assert p.returncode == 0, p.stderr
    # the weights file keeps the reference's name when scanning (plmDCA.hpp:550)
    assert (tmp_path / "first.weights").exists() and not (tmp_path / "first_scan.weights").exists()
    lines = (tmp_path / "first_scan.SuperDCA_couplings.1-based.all").read_text().splitlines()
    st1, tr1 = sd.prepare_alignment(a)
    st2, tr2 = sd.prepare_alignment(b)
    L1, L2 = st1.shape[1], st2.shape[1]
    assert len(lines) == L1 * L2
    with sd.Context(st2) as ctx:
        w2, _ = ctx.reweight(0.9)
    with sd.Context(np.ascontiguousarray(np.concatenate([st1, st2], axis=1))) as ctx:
        ctx.set_weights(w2)
        ctx.set_scan_split(L1)
        rows, _, _, _ = ctx.solve(np.arange(L1, dtype=np.int32))
    sd.write_scan_couplings(str(tmp_path / "direct.all"), np.ascontiguousarray(rows[:, L1:]), tr1, tr2, 1)
    assert_same_couplings(lines, (tmp_path / "direct.all").read_text().splitlines())


def test_cli_scan_against_a_one_locus_second_alignment(tmp_path):
    # any non-empty second alignment is scanned (plmDCA.hpp:491), one locus included
    d = synth.generate(B=120, N=30, clades=6, pairs=3, seed=607)
    a, b = d["chars"][:, :20], d["chars"][:, 20:]
    _, kept = sd.prepare_alignment(b)
    one = np.ascontiguousarray(b[:, kept[0]:kept[0] + 1])  # a column that passes the filter
    for name, part in (("first.fa", a), ("second.fa", one)):
        synth.write_fasta(str(tmp_path / name), part)
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "superdca_amd", "sdca")
    p = subprocess.run([exe, "first.fa", "second.fa"], capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert p.returncode == 0, p.stderr
    L1 = sd.prepare_alignment(a)[0].shape[1]
    lines = (tmp_path / "first_scan.SuperDCA_couplings.1-based.all").read_text().splitlines()
    assert len(lines) == L1
    assert all(t.split()[2] == "1" for t in lines)


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
