"""CPU tests of the host side of libsdca.so: ABI surface, alignment pipeline, writer, CLI."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import superdca_amd as sd
from superdca_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sdca.h")).read()
    return sorted(set(re.findall(r"\b(sdca_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(sd.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 12
    for name in syms:
        assert hasattr(lib, name), name
    assert sorted(sd.EXPORTED) == syms


def test_default_opts_match_reference_defaults():
    o = sd.default_opts()
    assert o.grad_threshold == 1e-3  # plmDCA_options.cpp:61
    assert o.max_iter == 2000        # plmDCA.hpp:219
    assert o.history == 0  # the formulation's default: 32 Gram, 10 primal
    assert o.precision == sd.PREC_FP32
    assert o.coarse_factor == 4  # split precision: BF16 high parts alone while ||g||_inf >= 4 * threshold
    assert o.j_threshold == 0.0  # SuperDCA's stopping rule alone unless a coupling threshold is asked for
    assert o.fine_factor == 10
    assert o.subspace_threshold == 3e-2


@pytest.mark.parametrize("bad", [{"coarse_factor": -1}, {"pipeline_groups": 3}, {"history": 7},
                                 {"gemm_tn_energy": 192}, {"gemm_gm_grad": 65}, {"precision": 9},
                                 {"fine_factor": -1}, {"j_threshold": -1e-6}, {"subspace_rank": 64},
                                 {"subspace_rank": 128, "subspace_threshold": 0.0}])
def test_invalid_options_are_rejected_before_any_device_call(bad):
    # option checks precede device selection, so they hold without a GPU
    st = np.zeros((4, 3), dtype=np.uint8)
    with pytest.raises(sd.SdcaError) as e:
        sd.Context(st, **bad)
    assert e.value.code == sd.SDCA_E_ARG, bad


def test_prepare_alignment_matches_column_stats():
    d = synth.generate(B=120, N=400, clades=6, pairs=5, seed=21)
    chars = d["chars"].copy()
    rng = np.random.default_rng(0)
    # make some loci fail: monomorphic, too gappy, too rare
    chars[:, 10] = ord("A")
    chars[: int(0.2 * 120), 11] = ord("-")
    chars[:, 12] = ord("C")
    chars[0, 12] = ord("G")  # maf 1/120 < 1 %
    chars[:, 13] = np.where(rng.random(120) < 0.5, ord("a"), ord("t"))  # lower case counts as alleles
    states, loci = sd.prepare_alignment(chars)
    keep = synth.passes_filter(chars)
    np.testing.assert_array_equal(loci, np.flatnonzero(keep))
    assert 10 not in loci and 11 not in loci and 12 not in loci and 13 in loci
    st2, loci2 = sd.prepare_alignment(chars, no_filter=True)
    assert loci2.size == chars.shape[1]


def test_envelope_state_order():
    # alleles by descending count, ties A<C<G<T, gap = q-1; four alleles dropped at q=4
    col = np.array([b"T", b"T", b"T", b"G", b"G", b"C", b"C", b"-", b"N", b"T"]).view(np.uint8)
    four = np.array([b"A", b"C", b"G", b"T", b"A", b"C", b"G", b"T", b"A", b"A"]).view(np.uint8)
    chars = np.stack([col, four], axis=1)
    states, loci = sd.prepare_alignment(chars, maf_threshold=0.0, gap_threshold=1.0)
    np.testing.assert_array_equal(loci, [0])
    # T (4) -> 0, C and G tie (2 each): C < G -> C=1, G=2, gaps and N -> 3
    np.testing.assert_array_equal(states[:, 0], [0, 0, 0, 2, 2, 1, 1, 3, 3, 0])
    states5, loci5 = sd.prepare_alignment(chars, maf_threshold=0.0, gap_threshold=1.0, q=5)
    np.testing.assert_array_equal(loci5, [0, 1])
    np.testing.assert_array_equal(states5[:, 1], [0, 1, 2, 3, 0, 1, 2, 3, 0, 0])
    assert states5[7, 0] == 4


def test_writer_matches_golden(tmp_path):
    S = np.load(os.path.join(GOLD, "couplings_tiny_S.npy"))
    out = tmp_path / "c.all"
    sd.write_couplings(str(out), S, np.array([2, 5, 7, 11], dtype=np.int32), 1)
    assert out.read_text() == open(os.path.join(GOLD, "couplings_tiny.all")).read()


def test_writer_rounding_matches_printf(tmp_path):
    rng = np.random.default_rng(3)
    L = 300
    S = (rng.random((L, L)) ** 3 * rng.choice([0.01, 0.1, 1.0, 7.0], (L, L))).astype(np.float32)
    S[rng.random((L, L)) < 0.01] = 0.0
    tr = np.sort(rng.choice(100000, L, replace=False)).astype(np.int32)
    out = tmp_path / "r.all"
    sd.write_couplings(str(out), S, tr, 1)
    lines = out.read_text().splitlines()
    assert len(lines) == L * (L - 1) // 2
    k = 0
    for r in range(L):
        for n in range(r):
            v = float(np.float32(S[r, n] + S[n, r])) * 0.5
            assert lines[k] == "%.8f %d %d" % (v, tr[r] + 1, tr[n] + 1)
            k += 1


def _full_file_values(S):
    # the symmetrised value of every pair, in the order of the full file (plmDCA.hpp:684-691)
    L = S.shape[0]
    vals = []
    for r in range(L):
        for n in range(r):
            vals.append(float(np.float32(S[r, n] + S[n, r])) * 0.5)
    return vals


@pytest.mark.parametrize("top_n,min_score", [(1, float("-inf")), (25, float("-inf")), (0, 0.5), (40, 0.9),
                                             (10 ** 9, float("-inf"))])
def test_top_writer_lists_the_largest_lines_of_the_full_file(tmp_path, top_n, min_score):
    rng = np.random.default_rng(8)
    L = 120
    S = rng.random((L, L)).astype(np.float32)
    S[rng.random((L, L)) < 0.02] = np.nan  # never listed
    S[:10, :10] = 0.25  # equal values: file order
    tr = np.sort(rng.choice(50000, L, replace=False)).astype(np.int32)
    sd.write_couplings(str(tmp_path / "f.all"), S, tr, 0)
    full = (tmp_path / "f.all").read_text().splitlines()
    vals = _full_file_values(S)
    assert len(full) == len(vals)
    order = sorted((k for k, v in enumerate(vals) if v >= min_score), key=lambda k: (-vals[k], k))
    if top_n:
        order = order[:top_n]
    sd.write_couplings_top(str(tmp_path / "f.top"), S, tr, 0, top_n if top_n < 10 ** 9 else 0, min_score)
    assert (tmp_path / "f.top").read_text().splitlines() == [full[k] for k in order]


def test_top_writer_rejects_bad_arguments(tmp_path):
    S = np.zeros((3, 3), dtype=np.float32)
    tr = np.arange(3, dtype=np.int32)
    with pytest.raises(sd.SdcaError, match="top_n < 0"):
        sd.write_couplings_top(str(tmp_path / "x.top"), S, tr, 1, -1)
    with pytest.raises(sd.SdcaError, match="NaN min_score"):
        sd.write_couplings_top(str(tmp_path / "x.top"), S, tr, 1, 5, float("nan"))
    with pytest.raises(sd.SdcaError, match="cannot open"):
        sd.write_couplings_top(str(tmp_path / "no_such_dir" / "x.top"), S, tr, 1, 5)
    with pytest.raises(sd.SdcaError, match="cannot open"):
        sd.write_couplings(str(tmp_path / "no_such_dir" / "x.all"), S, tr, 1)


def test_filter_reads_either_case_and_treats_other_letters_as_gaps():
    # lower case is the same allele; N and IUPAC codes count as gaps (MAF over non-gaps)
    col_upper = np.frombuffer(b"AAAAAAAACC", np.uint8)
    col_lower = np.frombuffer(b"aaaaAAAAcC", np.uint8)
    col_n = np.frombuffer(b"AAAAAAACCN", np.uint8)  # 1 gap of 10
    col_iupac = np.frombuffer(b"AAAARYCCAA", np.uint8)  # 2 gaps of 10 > 0.15: dropped
    ch = np.stack([col_upper, col_lower, col_n, col_iupac], axis=1)
    st, kept = sd.prepare_alignment(ch)
    assert kept.tolist() == [0, 1, 2]
    np.testing.assert_array_equal(st[:, 0], st[:, 1])
    assert st[9, 2] == 3 and st[:, 2].tolist()[:7] == [0] * 7  # N is the gap state (q - 1)


def test_scan_writer_lines(tmp_path):
    # plmDCA.hpp:672-681: every first-alignment locus against every second-alignment locus,
    # the float score printed with 8 fixed digits, no symmetrisation
    rng = np.random.default_rng(12)
    K, L2 = 7, 33
    S = (rng.random((K, L2)) ** 2 * 3).astype(np.float32)
    t1 = np.sort(rng.choice(1000, K, replace=False)).astype(np.int32)
    t2 = np.sort(rng.choice(5000, L2, replace=False)).astype(np.int32)
    sd.write_scan_couplings(str(tmp_path / "s.all"), S, t1, t2, 1)
    lines = (tmp_path / "s.all").read_text().splitlines()
    assert lines == ["%.8f %d %d" % (float(S[k, n]), t1[k] + 1, t2[n] + 1) for k in range(K) for n in range(L2)]
    with pytest.raises(ValueError):
        sd.write_scan_couplings(str(tmp_path / "x.all"), S, t1[:3], t2, 1)


def test_uneven_batch_shapes():
    cfgs = synth.c3_workloads()
    assert [c["seed"] for c in cfgs] == list(range(3000, 3048))
    for c in cfgs:
        assert 100 <= c["B"] <= 3000 and 500 <= c["N"] <= 20000 and 5 <= c["clades"] <= 50
        assert c["pairs"] == c["N"] // 100
    assert cfgs == synth.c3_workloads()  # deterministic
    assert len({(c["B"], c["N"]) for c in cfgs}) == 48
    small = min(cfgs, key=lambda c: c["B"] * c["N"])
    d = synth.generate(**small)
    assert d["chars"].shape == (small["B"], small["N"]) and synth.passes_filter(d["chars"]).all()


def test_generator_is_deterministic_and_filter_clean():
    a = synth.generate(B=80, N=300, clades=5, pairs=4, seed=5)
    b = synth.generate(B=80, N=300, clades=5, pairs=4, seed=5)
    np.testing.assert_array_equal(a["chars"], b["chars"])
    assert synth.passes_filter(a["chars"]).all()
    assert a["pairs"].shape == (4, 2)
    i, j = a["pairs"][0]
    # the planted partner copies a fixed map of locus i in most sequences
    ci, cj = a["chars"][:, i], a["chars"][:, j]
    agree = max(np.mean([(cj[ci == v] == cj[ci == v][0]).mean() for v in np.unique(ci) if (ci == v).sum() > 3]), 0)
    assert agree > 0.7


def test_cli_help_and_errors(tmp_path):
    exe = os.path.join(ROOT, "superdca_amd", "sdca")
    if not os.path.exists(exe):
        from superdca_amd import build
        build.build_cli()
    p = subprocess.run([exe, "--help"], capture_output=True, text=True)
    assert p.returncode == 0 and "--gradient-threshold" in p.stdout and "--maf-threshold" in p.stdout
    assert "--output-top" in p.stdout and "--output-min-score" in p.stdout
    p = subprocess.run([exe], capture_output=True, text=True)
    assert p.returncode != 0 and "No alignment file" in p.stderr
    p = subprocess.run([exe, str(tmp_path / "missing.fasta")], capture_output=True, text=True)
    assert p.returncode != 0 and "Could not get alignment" in p.stderr
    p = subprocess.run([exe, "--bogus-flag", "x.fa"], capture_output=True, text=True)
    assert p.returncode != 0
    fa = tmp_path / "t.fa"
    fa.write_text(">a\nACGT\n>b\nACG\n")
    p = subprocess.run([exe, str(fa)], capture_output=True, text=True)
    assert p.returncode != 0 and "different lengths" in p.stderr
    fa.write_text("ACGT\n")
    p = subprocess.run([exe, str(fa)], capture_output=True, text=True)
    assert p.returncode != 0 and "not in FASTA format" in p.stderr
    fa.write_text(">a\nACGT\n>b\nAGGT\n")
    (tmp_path / "cols.txt").write_text("1,9\n")
    p = subprocess.run([exe, "--filterlistfile", str(tmp_path / "cols.txt"), str(fa)], capture_output=True, text=True)
    assert p.returncode != 0 and "filter list index out of range" in p.stderr
    # --norm-of-mean-scoring writes its own six columns: the top-score output does not apply
    p = subprocess.run([exe, "--norm-of-mean-scoring", "--output-top", "10", str(fa)], capture_output=True, text=True)
    assert p.returncode != 0 and "six-column" in p.stderr
    p = subprocess.run([exe, "--chunk-loci", "-3", str(fa)], capture_output=True, text=True)
    assert p.returncode != 0 and "--chunk-loci" in p.stderr
    p = subprocess.run([exe, "--help"], capture_output=True, text=True)
    assert "--norm-of-mean-scoring" in p.stdout and "--coupling-gradient-threshold" in p.stdout


def test_context_without_gpu_fails_loudly():
    import torch  # noqa: F401  (only to learn whether a device is visible)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(sd.SdcaError):
        sd.Context(np.zeros((4, 3), dtype=np.uint8))


def test_bench_refuses_more_than_one_gpu():
    import sys

    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2 and "one GPU" in (p.stdout + p.stderr)


def _cli():
    exe = os.path.join(ROOT, "superdca_amd", "sdca")
    if not os.path.exists(exe):
        from superdca_amd import build
        build.build_cli()
    return exe


def _toy_alignment():
    # 50 loci that pass the default filter plus a monomorphic column (input column 10)
    # and a half-gap column (input column 31) that do not
    d = synth.generate(B=40, N=50, clades=4, pairs=2, seed=9)
    ch = d["chars"]
    mono = np.full((40, 1), ord("A"), np.uint8)
    gappy = np.where(np.arange(40)[:, None] % 2 == 0, ord("-"), ord("C")).astype(np.uint8)
    return np.concatenate([ch[:, :10], mono, ch[:, 10:30], gappy, ch[:, 30:]], axis=1)


def _write_wrapped_fasta(path, chars, names):
    # each sequence over two lines; every other record with CRLF line ends
    with open(path, "wb") as fh:
        for b, nm in enumerate(names):
            eol = b"\r\n" if b % 2 else b"\n"
            s = chars[b].tobytes()
            fh.write(b">" + nm.encode() + eol + s[:20] + eol + s[20:] + eol)


def _read_fasta(path):
    names, seqs = [], []
    for line in open(path).read().splitlines():
        if line.startswith(">"):
            names.append(line[1:])
            seqs.append("")
        elif line:
            seqs[-1] += line
    return names, seqs


def test_cli_writes_the_filtered_alignment_and_loci_list(tmp_path):
    # the host front end only: FASTA in, filter, the .fasta/.loci pair out (written before
    # any device is touched; the run may then stop for want of a GPU)
    ch = _toy_alignment()
    names = [f"s{b}" for b in range(ch.shape[0])]
    _write_wrapped_fasta(tmp_path / "toy.fa", ch, names)
    subprocess.run([_cli(), "--output-filtered-alignment", "--no-coupling-output", "toy.fa"], cwd=tmp_path,
                   capture_output=True, text=True, timeout=300)
    _, kept = sd.prepare_alignment(ch)
    assert 10 not in kept and 31 not in kept and kept.size == 50
    got_names, got = _read_fasta(tmp_path / "toy.fasta")
    assert got_names == names
    assert got == [ch[b, kept].tobytes().decode() for b in range(ch.shape[0])]
    loci = (tmp_path / "toy.loci").read_text().strip().split(",")
    assert [int(v) for v in loci] == (kept + 1).tolist()  # output indexing base 1


def test_cli_sample_and_column_lists(tmp_path):
    ch = _toy_alignment()
    names = [f"s{b}" for b in range(ch.shape[0])]
    _write_wrapped_fasta(tmp_path / "toy.fa", ch, names)
    keep_names = names[::2]
    (tmp_path / "samples.txt").write_text("\n".join(">" + n if i % 3 == 0 else n for i, n in enumerate(keep_names)) + "\n")
    cols = list(range(0, 52, 2)) + [10, 31]  # 1-based list below; duplicates removed, sorted
    (tmp_path / "cols.txt").write_text(",".join(str(c + 1) for c in cols[:10]) + "\n" +
                                       " ".join(str(c + 1) for c in cols[10:]) + "\n")
    subprocess.run([_cli(), "--samplelistfile", "samples.txt", "--filterlistfile", "cols.txt",
                    "--input-indexing-base", "1", "--output-filtered-alignment", "--no-coupling-output", "toy.fa"],
                   cwd=tmp_path, capture_output=True, text=True, timeout=300)
    rows = list(range(0, 40, 2))
    sel = sorted(set(cols))
    sub = ch[np.ix_(rows, sel)]
    _, kept = sd.prepare_alignment(sub)
    got_names, got = _read_fasta(tmp_path / "toy.fasta")
    assert got_names == keep_names
    assert got == [sub[i, kept].tobytes().decode() for i in range(len(rows))]
    loci = [int(v) for v in (tmp_path / "toy.loci").read_text().strip().split(",")]
    assert loci == [sel[k] + 1 for k in kept]  # original input columns, base 1


C_PROGRAM = r"""
#include <stdio.h>
#include <string.h>
#include "sdca.h"

int main(void)
{
    sdca_opts o;
    sdca_default_opts(&o);
    /* 3 sequences x 4 columns: columns 1 and 3 polymorphic, 0 monomorphic, 2 all gaps */
    const char chars[] = "AC-G" "AT-G" "AC-T";
    uint8_t states[12];
    int32_t loci[4], L = 0;
    int rc = sdca_prepare_alignment(chars, 3, 4, 0.01, 0.15, 0, 4, states, loci, &L);
    printf("threshold %g max_iter %d rc %d L %d loci", o.grad_threshold, o.max_iter, rc, L);
    for (int k = 0; k < L; ++k) printf(" %d", loci[k]);
    printf("\n");
    sdca_ctx *ctx = NULL;
    rc = sdca_ctx_create(&ctx, states, 3, L, 4, 3, &o);
    if (rc == SDCA_OK) {
        printf("context created\n");
        sdca_ctx_destroy(ctx);
    } else {
        printf("context refused: %d %s\n", rc, strlen(sdca_last_error()) ? "with message" : "without message");
    }
    rc = sdca_prepare_alignment(NULL, 3, 4, 0.01, 0.15, 0, 4, states, loci, &L);
    printf("null input rc %d\n", rc);
    return 0;
}
"""


def test_cli_two_alignment_argument_checks(tmp_path):
    exe = _cli()
    a = tmp_path / "a.fa"
    b = tmp_path / "b.fa"
    a.write_text(">s0\nACGT\n>s1\nAGGT\n")
    b.write_text(">s0\nACGT\n")
    p = subprocess.run([exe, str(a), str(b), str(a)], capture_output=True, text=True)
    assert p.returncode != 0 and "more than 2 alignments" in p.stderr
    p = subprocess.run([exe, str(a), str(b)], capture_output=True, text=True)
    assert p.returncode != 0 and "different numbers of sequences" in p.stderr
    (tmp_path / "cols.txt").write_text("1\n")
    p = subprocess.run([exe, "--filterlistfile", str(tmp_path / "cols.txt"), str(a), str(a)], capture_output=True,
                       text=True)
    assert p.returncode != 0 and "not supported with two alignments" in p.stderr
    p = subprocess.run([exe, "--output-filtered-alignment", str(a), str(a)], capture_output=True, text=True)
    assert p.returncode != 0 and "not supported with two alignments" in p.stderr


def test_plain_c_program_uses_the_abi(tmp_path):
    # the boundary is plain C: the header compiles as C99 and a C program links libsdca.so
    # (host entry points only; context creation reports a status when no GPU is present)
    import shutil

    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not found")
    src = tmp_path / "use_abi.c"
    src.write_text(C_PROGRAM)
    exe = str(tmp_path / "use_abi")
    libdir = os.path.dirname(sd.LIB_PATH)
    subprocess.run([gcc, "-std=c99", "-Wall", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"), str(src),
                    "-o", exe, "-L", libdir, "-lsdca", "-Wl,-rpath," + libdir], check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120).stdout.splitlines()
    assert out[0] == "threshold 0.001 max_iter 2000 rc 0 L 2 loci 1 3"
    assert out[1] == "context created" or out[1] == f"context refused: {sd.SDCA_E_HIP} with message"
    assert out[2] == f"null input rc {sd.SDCA_E_ARG}"


def test_bench_roofline_fields_from_timings():
    # the benchmark's roofline object from a library timing record (no device needed):
    # the dominant kernel by device time, achieved = algorithmic work / time, frac against
    # the MI355X peak, traffic = algorithmic bytes per launch x the measured counter ratio
    import importlib.util
    import sys

    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    argv = sys.argv
    try:
        sys.argv = ["bench.py"]
        spec.loader.exec_module(bench)
    finally:
        sys.argv = argv
    r = 1000
    tm = dict(gemm_energy_ms=2.2 * r, flops_energy=1.6e15, n_energy_launches=r,
              gemm_grad_ms=2.8 * r, flops_grad=1.6e15, n_grad_launches=r,
              update_ms=3.1 * r, bytes_update=15.6e9 * r, dots_ms=2.1 * r, bytes_dots=12.5e9 * r, n_evals=r,
              softmax_ms=0.1 * r, score_ms=0.05 * r)
    # the primal formulation (split BF16)
    rl = bench.roofline(tm, False, True)
    assert rl["kernel"] == "k_update" and rl["bound"] == "hbm" and rl["unit"] == "GB/s"
    assert rl["achieved"] == pytest.approx(15.6e9 / 3.1e-3 / 1e9, rel=1e-4)
    assert rl["frac"] == pytest.approx(rl["achieved"] / 8000.0, rel=1e-4)
    assert rl["traffic"] == pytest.approx(15.6e9 * 13.34 / 13.10, rel=1e-6)
    assert rl["launches"] == r and rl["avg_launch_us"] == pytest.approx(3100.0)
    phases, weighted = bench.phase_fractions(tm, False, True)
    assert set(phases) == {"k_gemm_nt<EpiEnergy>", "k_gemm_nt<EpiGrad>", "k_update", "k_dots"}
    assert phases["k_gemm_nt<EpiEnergy>"]["rate"] == pytest.approx(1.6e15 / 2.2 / 1e12, rel=1e-3)
    assert phases["k_gemm_nt<EpiGrad>"]["frac"] == pytest.approx(1.6e15 / 2.8 / 1e12 / 2500.0, abs=1e-4)
    fr = [p["frac"] for p in phases.values()]
    ms = [p["ms"] for p in phases.values()]
    assert weighted == pytest.approx(sum(f * m for f, m in zip(fr, ms)) / sum(ms), abs=1e-3)
    # the Gram formulation: the FP64 GEMM against the FP64 MFMA peak, traffic per flop
    tg = dict(tm, gemm_energy_ms=0.0, flops_energy=0.0, n_energy_launches=0, gemm_grad_ms=4.0 * r,
              flops_grad=2.0e11 * r, gstep_ms=1.0 * r, bytes_gstep=0.5 * tm["bytes_update"])
    rl = bench.roofline(tg, True, True)
    assert rl["kernel"] == "k_dgemm" and rl["bound"] == "mfma" and rl["peak"] == 78.6
    assert rl["traffic"] == pytest.approx(bench.HBM_BYTES_PER_FLOP["k_dgemm"] * 2.0e11, rel=1e-6)
    assert rl["achieved"] == pytest.approx(2.0e11 / 4.0e-3 / 1e12, rel=1e-4)
    assert rl["frac"] == pytest.approx(50.0 / 78.6, rel=1e-4)
    phases, _ = bench.phase_fractions(tg, True, True)
    assert set(phases) == {"k_dgemm", "k_gstep", "k_gsoftmax", "k_gdots"}
    # with the subspace warm start: its rounds as a phase of their own; the roofline names
    # the single kernel (k_dgemm or k_gstep) with the larger device time, never a group
    ts = dict(tg, subspace_ms=1.0 * r, flops_sub=1.0e10 * r, n_sub_evals=r, update_ms=9.0 * r)
    phases, _ = bench.phase_fractions(ts, True, True)
    assert set(phases) == {"k_dgemm", "k_gstep", "k_gsoftmax", "k_gdots", "subspace_rounds"}
    assert phases["subspace_rounds"]["rate"] == pytest.approx(1.0e10 / 1.0e-3 / 1e12, rel=1e-3)
    assert bench.roofline(ts, True, True)["kernel"] == "k_dgemm"
    tk = dict(ts, gstep_ms=6.0 * r)
    rl = bench.roofline(tk, True, True)
    assert rl["kernel"] == "k_gstep" and rl["bound"] == "hbm" and rl["peak"] == 8000.0
    assert rl["achieved"] == pytest.approx(0.5 * tm["bytes_update"] / (6.0 * r * 1e-3) / 1e9, rel=1e-4)


def test_every_hip_compile_disables_packed_fp32_and_targets_gfx950():
    # packed FP32 instructions are kept out of every kernel build
    from superdca_amd import build

    flags = build.HIP_FLAGS
    assert "--offload-arch=gfx950" in flags
    k = flags.index("-packed-fp32-ops")
    assert flags[k - 3:k] == ["-Xclang", "-target-feature", "-Xclang"]


def test_noalloc_reweight_refuses_bad_arguments_before_any_device_call():
    # null buffers, shapes outside the envelope and alignments beyond the 64 MiB scratch are
    # refused before a device is touched (the addresses below are never dereferenced)
    fake = 4096
    cases = [(None, 10, 10, 4, fake), (fake, 0, 10, 4, fake), (fake, 10, 10, 9, fake), (fake, 10, 10, 4, None),
             (fake, 3000, 94000, 5, fake)]
    for d_states, B, L, q, d_w in cases:
        with pytest.raises(sd.SdcaError) as e:
            sd.reweight_noalloc(d_states, B, L, q, 0.9, None, None, d_w)
        assert e.value.code == sd.SDCA_E_ARG, (B, L, q)
