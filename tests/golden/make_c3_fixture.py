"""Writes tests/golden/c3_fixture.npz (run from the repository root; a few minutes on eight
cores).

The uneven batch (superdca_amd.synth.c3_workloads: 48 alignments, seeds 3000-3047), alignment
k prepared with the state envelope q_env = 3 + k % 3.  For each, 4 loci drawn at random
(seed [seed, 1]) are solved by the FP64 oracle to its optimum (||g||_inf < 1e-8, or below
1e-6 where its line search stops at the FP64 rounding floor) and scored with the reference
gauge and Frobenius norm (plmDCA.hpp:366-394).  Kept per alignment k:

    q_k, L_k, w_sum_k     the envelope, the filtered loci, the sum of the weights
    loci_k, f_k           the 4 loci and their optimal objective values
    top_idx_k, top_val_k  the 200 largest entries of each locus' score row
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

NLOCI, TOP = 4, 200


def envelope(k: int) -> int:
    return 3 + k % 3


def main() -> None:
    import oracle as orc
    import superdca_amd as sd
    from superdca_amd import synth

    nt = os.cpu_count() or 8
    out = {}
    t0 = time.time()
    for k, cfg in enumerate(synth.c3_workloads()):
        q = envelope(k)
        st, _ = sd.prepare_alignment(synth.generate(**cfg)["chars"], q=q)
        B, L = st.shape
        w, _ = orc.weights(st, 0.9, nthreads=nt)
        _, lh, lJ = orc.lambdas(w)
        loci = np.sort(np.random.default_rng([cfg["seed"], 1]).choice(L, NLOCI, replace=False)).astype(np.int32)
        rows, _, stats = orc.solve_loci(st, w, lh, lJ, loci, q=q, gtol=1e-8, max_iter=20000, nthreads=nt)
        ok = (stats["status"] == 0) | ((stats["status"] == 2) & (stats["gnorm"] < 1e-6))
        assert np.all(ok), (k, stats["gnorm"][~ok])
        top = np.argsort(-rows, axis=1, kind="stable")[:, :TOP].astype(np.int32)
        out.update({f"q_{k}": q, f"L_{k}": L, f"w_sum_{k}": float(w.sum()), f"loci_{k}": loci,
                    f"f_{k}": stats["fval"], f"top_idx_{k}": top, f"top_val_{k}": np.take_along_axis(rows, top, 1)})
        print(f"alignment {k} (seed {cfg['seed']}): B={B} L={L} q={q}, max ||g|| {stats['gnorm'].max():.1e}, "
              f"{time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "c3_fixture.npz"), n=len(synth.c3_workloads()), **out)


if __name__ == "__main__":
    main()
