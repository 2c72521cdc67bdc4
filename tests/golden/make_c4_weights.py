"""Writes tests/golden/c4_weights.npz (run from the repository root; about 20 s on
eight cores).

The C4 workload (B = 3000 sequences, L = 94,000 loci, q_env = 5, seed 1004 of
superdca_amd.synth) reweighted by the oracle (calculate_weights, plmDCA_optimizer_parameters.hpp:39-98)
twice: at the default identity threshold 0.9 with multiplicity 1, and at threshold
0.9805 with multiplicities drawn from {1, 2, 3, 4} (seed 44).  Sequences of one clade
share 0.980-0.981 of their positions, so the second threshold splits those pairs and the
comparison at plmDCA_optimizer_parameters.hpp:72 decides many of them by a few
mismatches.

    w, int_w              weights and neighbour counts (theta 0.9, multiplicity 1)
    theta_m, mult         the second threshold and the multiplicities
    w_m, int_w_m          the weights and counts with them
    B, L                  the alignment's shape
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

THETA_M = 0.9805


def main() -> None:
    import oracle as orc
    import superdca_amd as sd
    from superdca_amd import synth

    st, _ = sd.prepare_alignment(synth.workload("C4")["chars"], q=5)
    B, L = st.shape
    nt = os.cpu_count() or 8
    t = time.time()
    w, iw = orc.weights(st, 0.9, nthreads=nt)
    mult = np.random.default_rng(44).integers(1, 5, B).astype(np.uint32)
    w_m, iw_m = orc.weights(st, THETA_M, mult=mult, nthreads=nt)
    print(f"B={B} L={L}: two reweightings in {time.time() - t:.0f} s, B_eff {w.sum():.6f}, "
          f"neighbour counts {iw.min()}..{iw.max()}; at {THETA_M}: B_eff {w_m.sum():.6f}, "
          f"counts {iw_m.min()}..{iw_m.max()}", flush=True)
    np.savez_compressed(os.path.join(HERE, "c4_weights.npz"), w=w, int_w=iw, mult=mult, theta_m=THETA_M, w_m=w_m, int_w_m=iw_m,
                        B=B, L=L)


if __name__ == "__main__":
    main()
