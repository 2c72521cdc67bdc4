"""Writes tests/golden/c2_top10k.npz (run from the repository root; about 20 minutes on
eight cores).

The C2 workload (B = 1000 sequences, L = 5000 loci, q_env = 4, seed 1002 of
superdca_amd.synth) solved for every locus by the FP64 oracle to its optimum
(||g||_inf < 1e-8, or below 1e-6 where its line search stops at the FP64 rounding
floor), scored with the reference gauge and Frobenius norm, symmetrised as the
coupling writer does (double(float(S(r, n) + S(n, r))) * 0.5, plmDCA.hpp:667-693), and
reduced to what the GPU comparison needs:

    top_i, top_j, top_s   the 10,000 highest-scoring pairs (i < j) and their scores
    kth                   the 10,000-th highest score
    smp_i, smp_j, smp_s   5,000 further pairs drawn uniformly (seed 7) with their scores
    w_sum                 the sum of the sequence weights (ties the fixture to its input)
    gnorm_max             the largest final ||g||_inf of the oracle's loci
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def symmetrised(rows: np.ndarray) -> np.ndarray:
    return (rows + rows.T).astype(np.float32).astype(np.float64) * 0.5


def main() -> None:
    import oracle as orc
    import superdca_amd as sd
    from superdca_amd import synth

    d = synth.workload("C2")
    st, _ = sd.prepare_alignment(d["chars"])
    B, L = st.shape
    w, _ = orc.weights(st, 0.9, nthreads=os.cpu_count() or 8)
    _, lh, lJ = orc.lambdas(w)
    t = time.time()
    rows, _, stats = orc.solve_loci(st, w, lh, lJ, np.arange(L), gtol=1e-8, max_iter=20000,
                                    nthreads=os.cpu_count() or 8)
    ok = (stats["status"] == 0) | ((stats["status"] == 2) & (stats["gnorm"] < 1e-6))
    print(f"solved {L} loci in {time.time() - t:.0f} s, status counts {np.bincount(stats['status'])}, "
          f"max ||g||_inf {stats['gnorm'].max():.2e}", flush=True)
    assert np.all(ok), stats["gnorm"][~ok]
    S = symmetrised(rows)
    iu, ju = np.triu_indices(L, 1)
    sv = S[iu, ju]
    order = np.argsort(-sv, kind="stable")[:10000]
    rng = np.random.default_rng(7)
    smp = rng.choice(iu.size, 5000, replace=False)
    np.savez_compressed(os.path.join(HERE, "c2_top10k.npz"), top_i=iu[order].astype(np.int32),
                        top_j=ju[order].astype(np.int32), top_s=sv[order], kth=sv[order[-1]],
                        smp_i=iu[smp].astype(np.int32), smp_j=ju[smp].astype(np.int32), smp_s=sv[smp],
                        w_sum=float(w.sum()), B=B, L=L, gnorm_max=float(stats["gnorm"].max()))


if __name__ == "__main__":
    main()
