"""--norm-of-mean-scoring through the sdca CLI (GPU solve, file-backed block store, float
writer) against the FP64 oracle's optimum put through the oracle's restatement of
plmDCA.hpp:617-666 (one alignment) and :625-638 (two-alignment scan)."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import oracle as orc
import superdca_amd as sd
from superdca_amd import synth

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "superdca_amd", "sdca")


def run_cli(args, cwd):
    p = subprocess.run([EXE, *args], capture_output=True, text=True, cwd=cwd, timeout=300)
    assert p.returncode == 0, p.stderr
    return p


def test_norm_of_mean_cli_against_the_oracle(tmp_path):
    d = synth.generate(B=150, N=60, clades=6, pairs=4, seed=88)
    synth.write_fasta(str(tmp_path / "toy.fasta"), d["chars"])
    # chunks of 16 loci: the blocks reach the store from several solve calls
    run_cli(["--norm-of-mean-scoring", "--coupling-gradient-threshold", "3e-6", "--chunk-loci", "16", "toy.fasta"],
            tmp_path)
    lines = (tmp_path / "toy.SuperDCA_couplings.1-based.all").read_text().splitlines()
    st, tr = sd.prepare_alignment(d["chars"])
    B, L = st.shape
    assert len(lines) == L * (L - 1) // 2
    w, _ = orc.weights(st, 0.9, nthreads=8)
    _, lh, lJ = orc.lambdas(w)
    loci = np.arange(L, dtype=np.int32)
    _, X, so = orc.solve_loci(st, w, lh, lJ, loci, gtol=1e-9, max_iter=20000, nthreads=8, want_rows=False,
                              want_x=True)
    assert np.all((so["status"] == 0) | ((so["status"] == 2) & (so["gnorm"] < 1e-6)))
    got, want, k = [], [], 0
    for r in range(L):
        for n in range(r):
            t = lines[k].split()
            k += 1
            assert (int(t[1]), int(t[2])) == (tr[r] + 1, tr[n] + 1)
            jm, jij, jji = orc.nom_pair(orc.nom_block(X[r], L, 4, r, n), orc.nom_block(X[n], L, 4, n, r))
            want.append((jm, jij, jji, 0.5 * (jij + jji)))
            got.append(tuple(float(v) for v in (t[0], t[3], t[4], t[5])))
    got, want = np.array(got), np.array(want, dtype=np.float64)
    big = want[:, 0] > np.quantile(want[:, 0], 0.9)
    rel = np.abs(got - want)[big] / want[big]
    print(f"norm of mean: {len(lines)} pairs, top-decile rel err max {rel.max():.2e}, median {np.median(rel):.2e}")
    assert rel.max() < 1e-4


def test_norm_of_mean_scan_cli_against_the_oracle(tmp_path):
    d = synth.generate(B=120, N=40, clades=6, pairs=3, seed=89)
    a, b = d["chars"][:, :8], d["chars"][:, 8:]
    for name, part in (("first.fa", a), ("second.fa", b)):
        synth.write_fasta(str(tmp_path / name), part)
    run_cli(["--norm-of-mean-scoring", "--coupling-gradient-threshold", "3e-6", "first.fa", "second.fa"], tmp_path)
    lines = (tmp_path / "first_scan.SuperDCA_couplings.1-based.all").read_text().splitlines()
    st1, tr1 = sd.prepare_alignment(a)
    st2, tr2 = sd.prepare_alignment(b)
    L1, L2 = st1.shape[1], st2.shape[1]
    assert len(lines) == L1 * L2
    w2, _ = orc.weights(st2, 0.9)
    _, lh, lJ = orc.lambdas(w2)
    got, want = [], []
    for r in range(L1):
        # locus r fitted on the second alignment alone: the oracle's ordinary solve of locus 0
        # of the alignment [r, second]
        red = np.ascontiguousarray(np.concatenate([st1[:, r:r + 1], st2], axis=1))
        _, X, so = orc.solve_loci(red, w2, lh, lJ, np.array([0], dtype=np.int32), gtol=1e-9, max_iter=20000,
                                  want_rows=False, want_x=True)
        for u in range(L2):
            t = lines[r * L2 + u].split()
            assert (int(t[1]), int(t[2])) == (tr1[r] + 1, tr2[u] + 1)
            M = orc.nom_block(X[0], 1 + L2, 4, 0, 1 + u)
            want.append(float(orc.nom_pair(M, M)[1]))
            got.append(float(t[0]))
    got, want = np.array(got), np.array(want)
    big = want > np.quantile(want, 0.9)
    rel = np.abs(got - want)[big] / want[big]
    print(f"scan norm of mean: top-decile rel err max {rel.max():.2e}")
    assert rel.max() < 1e-4
