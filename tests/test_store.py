"""File-backed stores and the streaming writers (CPU; no GPU needed).

The streamed .all/.top files must equal, byte for byte, what the in-memory writers
produce from the whole score matrix; the --norm-of-mean-scoring writer must print, line
by line, the oracle's restatement of plmDCA.hpp:617-666 for the same solutions."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as orc
import superdca_amd as sd


@pytest.mark.parametrize("tile", [256, 0])
def test_streamed_couplings_equal_the_in_memory_writers(tmp_path, tile):
    L = 3000
    rng = np.random.default_rng(21)
    S = (rng.random((L, L), dtype=np.float32) * 0.1).astype(np.float32)
    S[5, 7] = S[7, 5] = np.float32(0.05)  # a tie with other pairs at the threshold below
    tr = rng.permutation(4 * L)[:L].astype(np.int32)
    sd.write_couplings(str(tmp_path / "mem.all"), S, tr, 1)
    sd.write_couplings_top(str(tmp_path / "mem.top"), S, tr, 1, top_n=5000, min_score=0.05)
    with sd.Store(str(tmp_path / "scores.store"), L, L) as st:
        # rows arrive in chunks of uneven size and in no particular order
        bounds = list(range(0, L, 137)) + [L]
        chunks = list(zip(bounds[:-1], bounds[1:]))
        for i in rng.permutation(len(chunks)):
            a, b = chunks[i]
            st.put_rows(a, S[a:b])
        st.write_couplings(tmp_path / "st.all", tmp_path / "st.top", tr, 1, top_n=5000, min_score=0.05, tile=tile)
    assert not (tmp_path / "scores.store").exists()  # the store's file is unlinked
    assert (tmp_path / "st.all").read_bytes() == (tmp_path / "mem.all").read_bytes()
    assert (tmp_path / "st.top").read_bytes() == (tmp_path / "mem.top").read_bytes()


def test_store_rejects_bad_arguments(tmp_path):
    with pytest.raises(sd.SdcaError):
        sd.Store(str(tmp_path / "a"), 0, 4)
    with sd.Store(str(tmp_path / "b"), 4, 4) as st:
        with pytest.raises(sd.SdcaError):
            st.put_rows(3, np.zeros((2, 4), dtype=np.float32))  # rows [3, 5) of 4
        with pytest.raises(sd.SdcaError):
            st.write_norm_of_mean(tmp_path / "x", np.arange(4), q=4)  # entry 1 is not a block store
    (tmp_path / "c").write_text("exists")
    with pytest.raises(sd.SdcaError):
        sd.Store(str(tmp_path / "c"), 2, 2)  # never overwrites a file


def nom_lines_oracle(x, L, q, loci, tr, base=1):
    """The six-column lines from the oracle's float restatement (orc_nom_*)."""
    lines = []
    for k, r in enumerate(loci):
        for u in range(k):
            n = loci[u]
            a = orc.nom_block(x[k], L, q, r, n)
            b = orc.nom_block(x[u], L, q, n, r)
            jm, jij, jji = (float(v) for v in orc.nom_pair(a, b))
            avg = float(np.float32(np.float32(jij) + np.float32(jji))) * 0.5
            lines.append(f"{jm:.8f} {tr[k] + base} {tr[u] + base} {jij:.8f} {jji:.8f} {avg:.8f}")
    return lines


@pytest.mark.parametrize("q,order", [(4, "ascending"), (5, "ascending"), (4, "shuffled")])
def test_norm_of_mean_writer_matches_the_oracle(tmp_path, q, order):
    L = 37  # three 16-locus blocks, the last one short
    rng = np.random.default_rng(q)
    loci = np.arange(L, dtype=np.int32) if order == "ascending" else rng.permutation(L).astype(np.int32)
    K = loci.size
    x = rng.normal(scale=0.05, size=(K, sd.nparams(L, q)))
    tr = (loci * 3 + 1).astype(np.int32)
    with sd.Store(str(tmp_path / "blocks.store"), K, K, q * q) as st:
        for k in rng.permutation(K):
            st.put_blocks(int(k), x[k], L, q, int(loci[k]), loci)
        st.write_norm_of_mean(tmp_path / "nom.all", tr, 1, q=q, tile=8)
    got = (tmp_path / "nom.all").read_text().splitlines()
    assert got == nom_lines_oracle(x, L, q, loci, tr)


def test_norm_of_mean_scan_writer_matches_the_oracle(tmp_path):
    L1, L2, q = 5, 23, 4
    L = L1 + L2
    rng = np.random.default_rng(9)
    loci = np.array([0, 3, 4], dtype=np.int32)
    x = rng.normal(scale=0.05, size=(loci.size, sd.nparams(L, q)))
    cols = np.arange(L1, L, dtype=np.int32)
    t1, t2 = np.array([10, 13, 14], dtype=np.int32), np.arange(L2, dtype=np.int32) + 100
    with sd.Store(str(tmp_path / "scan.store"), loci.size, L2, q * q) as st:
        for k, r in enumerate(loci):
            st.put_blocks(k, x[k], L, q, int(r), cols)
        st.write_norm_of_mean_scan(tmp_path / "scan.all", t1, t2, 1, q=q)
    want = []
    for k, r in enumerate(loci):
        for u, n in enumerate(cols):
            M = orc.nom_block(x[k], L, q, int(r), int(n))
            v = float(orc.nom_pair(M, M)[1])  # the norm of the gauged block alone
            want.append(f"{v:.8f} {t1[k] + 1} {t2[u] + 1}")
    assert (tmp_path / "scan.all").read_text().splitlines() == want
