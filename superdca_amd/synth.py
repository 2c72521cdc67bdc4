"""Seeded synthetic genome alignments for tests and benchmarks.

Model: B sequences over {A, C, G, T, -}.  Each sequence belongs to one of `clades`
clades and copies that clade's founder, with each site independently redrawn from the
locus' allele distribution at rate `mut` (2 %).  Each locus is biallelic with minor-allele
frequency ~ U(0.01, 0.5) and gap rate ~ U(0, 0.10); a fraction `tri_frac` (5 %) carries a
third allele.  Loci that fail the default filter (more than one allele, MAF >= 1 %,
gap <= 15 %) are redrawn, so all loci pass.  `pairs` planted couplings: for a pair (i, j)
locus j takes a fixed state mapping of locus i in each sequence with probability 0.8.

Named workloads (sizes of the reference runs this engine is measured on):
  C1  B=200,  N=1,000,  10 clades,   50 pairs, seed 1001
  C2  B=1000, N=5,000,  30 clades,  300 pairs, seed 1002
  C4  B=3000, N=94,000, 60 clades, 2000 pairs, seed 1004
"""
from __future__ import annotations

import numpy as np

NUC = np.frombuffer(b"ACGT", dtype=np.uint8)
GAP = ord("-")

WORKLOADS = {
    "C1": dict(B=200, N=1000, clades=10, pairs=50, seed=1001),
    "C2": dict(B=1000, N=5000, clades=30, pairs=300, seed=1002),
    "C4": dict(B=3000, N=94000, clades=60, pairs=2000, seed=1004),
}


def column_stats(chars: np.ndarray):
    """Per-locus (n_alleles, maf, gap_fraction, counts[4]) with the filter's definitions."""
    chars = np.asarray(chars, dtype=np.uint8)
    B = chars.shape[0]
    up = chars & 0xDF  # upper case
    counts = np.stack([(up == c).sum(axis=0) for c in NUC], axis=1)  # N x 4
    ngap = B - counts.sum(axis=1)
    nall = (counts > 0).sum(axis=1)
    srt = -np.sort(-counts, axis=1)
    nongap = B - ngap
    with np.errstate(divide="ignore", invalid="ignore"):
        maf = np.where((nall > 1) & (nongap > 0), srt[:, 1] / np.maximum(nongap, 1), 0.0)
    return nall, maf, ngap / B, counts


def passes_filter(chars, maf_threshold=0.01, gap_threshold=0.15):
    nall, maf, gapf, _ = column_stats(chars)
    return (nall > 1) & (maf >= maf_threshold) & (gapf <= gap_threshold)


def _categorical(rng, p, size):
    """Draws categories from per-column probability rows p (n x k) for `size` rows."""
    cp = np.cumsum(p, axis=1)
    cp[:, -1] = 1.0
    u = rng.random((size, p.shape[0]))
    return (u[:, :, None] > cp[None, :, :]).sum(axis=2).astype(np.int8)


def _draw_loci(rng, n, B, clade, nclades, mut, tri_frac, maf_range, gap_max):
    maf = rng.uniform(maf_range[0], maf_range[1], n)
    gap = rng.uniform(0.0, gap_max, n)
    tri = rng.random(n) < tri_frac
    third = np.where(tri, rng.uniform(0.2, 1.0, n) * np.minimum(maf, np.maximum(1.0 - 2.0 * maf, 0.0)), 0.0)
    nongap = 1.0 - gap
    p = np.stack([nongap * (1.0 - maf - third), nongap * maf, nongap * third, gap], axis=1)  # major, minor, third, gap
    founders = _categorical(rng, p, nclades)  # clades x n
    cats = founders[clade]  # B x n
    redraw = rng.random((B, n)) < mut
    fresh = _categorical(rng, p, B)
    cats = np.where(redraw, fresh, cats)
    perm = np.argsort(rng.random((n, 4)), axis=1)  # nucleotide of category k at each locus
    alle = NUC[perm]  # n x 4 ascii
    out = np.where(cats == 3, GAP, alle[np.arange(n)[None, :], np.minimum(cats, 2)]).astype(np.uint8)
    return out


def generate(B: int, N: int, clades: int, pairs: int, seed: int, mut: float = 0.02, tri_frac: float = 0.05,
             maf_range=(0.01, 0.5), gap_max: float = 0.10, maf_threshold: float = 0.01,
             gap_threshold: float = 0.15, chunk: int = 2048):
    """Returns {"chars": B x N uint8 ASCII, "pairs": P x 2 int (i, j), "clade": B}."""
    rng = np.random.default_rng(seed)
    clade = rng.integers(0, clades, B)
    chars = np.empty((B, N), dtype=np.uint8)
    for j0 in range(0, N, chunk):
        j1 = min(N, j0 + chunk)
        block = _draw_loci(rng, j1 - j0, B, clade, clades, mut, tri_frac, maf_range, gap_max)
        for _ in range(100):
            bad = np.flatnonzero(~passes_filter(block, maf_threshold, gap_threshold))
            if bad.size == 0:
                break
            block[:, bad] = _draw_loci(rng, bad.size, B, clade, clades, mut, tri_frac, maf_range, gap_max)
        else:
            raise RuntimeError("generator could not satisfy the filter")
        chars[:, j0:j1] = block
    # planted couplings
    planted = []
    order = rng.permutation(N)
    used = np.zeros(N, dtype=bool)
    cursor = 0
    for _ in range(pairs):
        while cursor < N and used[order[cursor]]:
            cursor += 1
        if cursor >= N:
            break
        i = int(order[cursor])
        used[i] = True
        cursor += 1
        for _attempt in range(20):
            while cursor < N and used[order[cursor]]:
                cursor += 1
            if cursor >= N:
                break
            j = int(order[cursor])
            cursor += 1
            col_i = chars[:, i]
            _, _, _, ci = column_stats(col_i[:, None])
            _, _, _, cj = column_stats(chars[:, j][:, None])
            ai = NUC[np.argsort(-ci[0], kind="stable")]
            aj = NUC[np.argsort(-cj[0], kind="stable")]
            bj = list(aj[:3])
            if cj[0][np.argsort(-cj[0], kind="stable")[2]] == 0:
                bj[2] = bj[1]
            if rng.random() < 0.5:
                bj[0], bj[1] = bj[1], bj[0]
            lut = np.full(256, GAP, dtype=np.uint8)
            for k in range(3):
                lut[ai[k]] = bj[k]
            lut[ai[3]] = bj[2]
            take = rng.random(B) < 0.8
            newcol = np.where(take, lut[col_i], chars[:, j]).astype(np.uint8)
            if passes_filter(newcol[:, None], maf_threshold, gap_threshold)[0]:
                chars[:, j] = newcol
                used[j] = True
                planted.append((i, j))
                break
    return {"chars": chars, "pairs": np.array(planted, dtype=np.int64).reshape(-1, 2), "clade": clade}


def c3_workloads():
    """The uneven batch: 48 independent alignments, seeds 3000-3047, each with
    B ~ U[100, 3000] sequences, N ~ U[500, 20000] loci, U[5, 50] clades and planted pairs
    1 % of N (the shapes are drawn from the seed's own stream [seed, 0])."""
    out = []
    for k in range(48):
        seed = 3000 + k
        r = np.random.default_rng([seed, 0])
        B = int(r.integers(100, 3001))
        N = int(r.integers(500, 20001))
        clades = int(r.integers(5, 51))
        out.append(dict(B=B, N=N, clades=clades, pairs=max(1, N // 100), seed=seed))
    return out


def workload(name: str, **override):
    cfg = dict(WORKLOADS[name])
    cfg.update(override)
    return generate(**cfg)


def write_fasta(path: str, chars: np.ndarray, prefix: str = "seq") -> None:
    with open(path, "wb") as fh:
        for b in range(chars.shape[0]):
            fh.write(b">%s%d\n" % (prefix.encode(), b))
            fh.write(chars[b].tobytes())
            fh.write(b"\n")
