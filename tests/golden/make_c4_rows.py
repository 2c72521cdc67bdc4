"""Writes tests/golden/c4_rows64.npz (run from the repository root; about 15 minutes on
eight cores).

The C4 workload (B = 3000 sequences, L = 94,000 loci, q_env = 5, seed 1004 of
superdca_amd.synth): 64 loci drawn at random (seed 64) solved by the FP64 oracle to its
optimum (||g||_inf < 1e-8, or below 1e-6 where its line search stops at the FP64
rounding floor) and scored with the reference gauge and Frobenius norm (score rows S(r, .),
plmDCA.hpp:366-394).  Kept per locus: the 200 largest entries of its row and 200 entries
drawn at random (seed 65), with their indices.

    loci                  the 64 loci
    top_idx, top_val      64 x 200 column indices and scores of the largest entries
    smp_idx, smp_val      64 x 200 random column indices (distinct from r) and scores
    w_sum, gnorm_max      the sum of the sequence weights; the largest final ||g||_inf
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main() -> None:
    import oracle as orc
    import superdca_amd as sd
    from superdca_amd import synth

    st, _ = sd.prepare_alignment(synth.workload("C4")["chars"], q=5)
    B, L = st.shape
    nt = os.cpu_count() or 8
    w, _ = orc.weights(st, 0.9, nthreads=nt)
    _, lh, lJ = orc.lambdas(w)
    loci = np.sort(np.random.default_rng(64).choice(L, 64, replace=False)).astype(np.int32)
    t = time.time()
    rows, _, stats = orc.solve_loci(st, w, lh, lJ, loci, q=5, gtol=1e-8, max_iter=20000, nthreads=nt)
    ok = (stats["status"] == 0) | ((stats["status"] == 2) & (stats["gnorm"] < 1e-6))
    print(f"solved {loci.size} loci in {time.time() - t:.0f} s, status counts {np.bincount(stats['status'])}, "
          f"max ||g||_inf {stats['gnorm'].max():.2e}", flush=True)
    assert np.all(ok), stats["gnorm"][~ok]
    top_idx = np.argsort(-rows, axis=1, kind="stable")[:, :200].astype(np.int32)
    top_val = np.take_along_axis(rows, top_idx, axis=1)
    rng = np.random.default_rng(65)
    smp_idx = np.empty((loci.size, 200), dtype=np.int32)
    for k, r in enumerate(loci):
        c = rng.choice(L - 1, 200, replace=False)
        smp_idx[k] = c + (c >= r)  # skip the locus itself
    smp_val = np.take_along_axis(rows, smp_idx, axis=1)
    np.savez_compressed(os.path.join(HERE, "c4_rows64.npz"), loci=loci, top_idx=top_idx, top_val=top_val,
                        smp_idx=smp_idx, smp_val=smp_val, w_sum=float(w.sum()),
                        gnorm_max=float(stats["gnorm"].max()), B=B, L=L)


if __name__ == "__main__":
    main()
