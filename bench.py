"""Throughput of the per-locus plmDCA solve on one MI355X: loci/s to convergence.

Workload: B = 3000 sequences x N = 94,000 loci (seeded synthetic genome alignment, 60
clades, 2 % mutation, 2000 planted pairs), q = 5 envelope, L-BFGS m = 10.  A step is a
batch of loci solved to convergence with the slots in flight refilled as loci finish.
Timing starts with the alignment, weights and lambda resident on the device (and, for the
Gram formulation, the alignment's identical-locus counts built once: reported) and ends
when the last score row is in host memory.

Headline (value): the Gram formulation -- each locus' couplings held as J = X_r^T c in
the span of the data, FP64 evaluations with one B x B x q GEMM per round -- warm-started
from the locus' solve restricted to the dominant 128-dimensional eigen-subspace of the
identical-locus counts (two B x 128 x q products per restricted evaluation), stopping at
||g||_2 < 1e-3 and ||g_J||_2 < 3e-5 (2-norms: bounds on the infinity norms of SuperDCA's
rule from above).  Its score error against the FP64 oracle is measured on the 64 loci of
the committed C4 fixture, solved with the headline settings.

Secondary legs on the same GPU:
  gram-cold  the same loci, the same rule, from a cold start (zero)
  gram-rule  the Gram formulation at ||g||_2 < 1e-3 alone
  primal     split-BF16 contractions over the K_x one-hot columns, ||g||_inf < 1e-3 and
             ||g_J||_inf < 3e-6, cold and warm-started from the network (trained
             beforehand on solutions of the C1 and C2 workloads, untimed) for the top-2048
             MI partners of every locus, on the same loci
  fp32       the plain path: primal FP32 MFMA contractions, cold start, the same rule
The MI partner table of the alignment is built once (reported, not in the loci/s).

Prints one JSON line: value (headline loci/s), the roofline of the headline's dominant
kernel (HIP events on the library's streams over the timed region; with two slot groups
on two streams a kernel's launch time includes the other group's kernels sharing the GPU,
so single-group rates are reported beside it), and a CPU baseline: the FP64 oracle
(block-compressed, one locus per thread) on a bounded sample of the same workload at
SuperDCA's rule, on the box's CPU share and on one thread.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"bf16_tflops": 2500.0, "fp32_tflops": 157.3, "fp64_tflops": 78.6, "hbm_gbs": 8000.0}
METRIC = "loci/s to converged plmDCA at B=3000,N=94k; warm vs cold start, same GPU"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def run_steps(ctx, order, pos, per_step, steps, label, solve, fvals=None, subs=None):
    """Times `steps` solve calls; returns (loci/s, iterations, evaluations, seconds, new pos).
    The final objective values are appended to `fvals` when given, the subspace warm
    start's evaluations to `subs`."""
    import superdca_amd as sd

    ctx.timing(reset=True)
    converged = 0
    iters, evals = [], []
    t0 = time.perf_counter()
    for k in range(steps):
        _, st, _, _ = solve(order[pos:pos + per_step])
        pos += per_step
        converged += int((st["status"] == sd.ST_CONVERGED).sum())
        iters.append(st["iter"])
        evals.append(st["nfeval"])
        if fvals is not None:
            fvals.append(st["fval"])
        if subs is not None and "nfeval_sub" in st:
            subs.append(st["nfeval_sub"])
        log(f"[bench] {label} step {k}: {converged} converged, {time.perf_counter() - t0:.2f}s")
    elapsed = time.perf_counter() - t0
    return converged / elapsed, np.concatenate(iters), np.concatenate(evals), elapsed, pos


# HBM bytes moved per algorithmic byte (memory-bound kernels) or per algorithmic flop (the
# GEMMs) of a kernel at C4, from rocprofv3 FETCH_SIZE and WRITE_SIZE passes scaled by a plain
# device copy (fetch x2.000, write x1.000); the profiles hold the commands and the
# per-dispatch tables.  k_gstep covers the step pass (k_gstep + k_gsoftmax).
HBM_PER_ALGORITHMIC_BYTE = {"k_update": 13.34 / 13.10, "k_dots": 10.51 / 10.50,
                            "k_gstep": 12.618 / 12.253, "k_gdots": 6.689 / 6.523}
HBM_BYTES_PER_FLOP = {"k_dgemm": 1.757e9 / 0.1134e12}


def kernel_groups(tm, gram, split):
    """(device ms, algorithmic work, bound, launches, peak) of each dominant kernel of a leg.
    Gram: k_dgemm is the B x B x q FP64 GEMM (one launch per round and group), the step pass
    (k_ghstep, k_gstep, k_gsoftmax) and the products pass (k_gdots, k_greduce) move the
    slot vectors.  Primal: the two split-BF16 (or FP32) contractions and the vector passes."""
    if gram:
        out = {
            "k_dgemm": (tm["gemm_grad_ms"], tm["flops_grad"], "mfma", tm["n_grad_launches"], PEAK["fp64_tflops"]),
            "k_gstep": (tm["gstep_ms"], tm["bytes_gstep"], "hbm", tm["n_grad_launches"], PEAK["hbm_gbs"]),
            # the rest of the step pass: k_ghstep, k_gsoftmax and the staged warm starts
            "k_gsoftmax": (tm["update_ms"] - tm["gstep_ms"], tm["bytes_update"] - tm["bytes_gstep"], "hbm",
                           tm["n_evals"], PEAK["hbm_gbs"]),
            "k_gdots": (tm["dots_ms"], tm["bytes_dots"], "hbm", tm["n_evals"], PEAK["hbm_gbs"]),
        }
        if tm.get("n_sub_evals"):
            # the subspace warm start's rounds (all of their kernels), against the FP64 MFMA peak
            out["subspace_rounds"] = (tm["subspace_ms"], tm["flops_sub"], "mfma", tm["n_sub_evals"],
                                      PEAK["fp64_tflops"])
        return out
    mp = PEAK["bf16_tflops"] if split else PEAK["fp32_tflops"]
    return {
        "k_gemm_nt<EpiGrad>": (tm["gemm_grad_ms"], tm["flops_grad"], "mfma", tm["n_grad_launches"], mp),
        "k_gemm_nt<EpiEnergy>": (tm["gemm_energy_ms"], tm["flops_energy"], "mfma", tm["n_energy_launches"], mp),
        "k_update": (tm["update_ms"], tm["bytes_update"], "hbm", tm["n_evals"], PEAK["hbm_gbs"]),
        "k_dots": (tm["dots_ms"], tm["bytes_dots"], "hbm", tm["n_evals"], PEAK["hbm_gbs"]),
    }


def roofline(tm, gram, split):
    """The dominant kernel's roofline: the largest by device time.  Gram: of the two single
    kernels timed on their own, k_dgemm and k_gstep (the other entries are groups of
    kernels).  Primal: of the four groups."""
    groups = kernel_groups(tm, gram, split)
    cands = ("k_dgemm", "k_gstep") if gram else tuple(groups)
    dom = max(cands, key=lambda k: groups[k][0])
    ms, work, bound, n, peak = groups[dom]
    scale, unit = (1e12, "TFLOP/s") if bound == "mfma" else (1e9, "GB/s")
    achieved = work / (ms * 1e-3) / scale
    if bound == "hbm":
        ratio = HBM_PER_ALGORITHMIC_BYTE.get(dom)
        basis = (f"HBM bytes per launch: algorithmic bytes x {ratio:.4f}, the measured (FETCH_SIZE + WRITE_SIZE) / "
                 "algorithmic ratio of this kernel") if ratio else None
    else:
        ratio = HBM_BYTES_PER_FLOP.get(dom)
        basis = (f"HBM bytes per launch: {ratio:.5f} bytes per algorithmic flop, the measured (FETCH_SIZE + "
                 "WRITE_SIZE) per flop of this kernel") if ratio else None
    traffic = round(ratio * work / max(n, 1)) if ratio else None
    return {"bound": bound, "kernel": dom, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
            "frac": round(achieved / peak, 5), "traffic": traffic, "traffic_basis": basis,
            "launches": int(n), "avg_launch_us": round(ms * 1e3 / max(n, 1), 1),
            "work_per_launch": round(work / max(n, 1) / scale, 4)}


def phase_fractions(tm, gram, split):
    """Roofline fraction of each dominant kernel (algorithmic work over its device time,
    against the MFMA or HBM peak) and their device-time-weighted mean."""
    out, weighted, total = {}, 0.0, 0.0
    for name, (ms, work, bound, n, peak) in kernel_groups(tm, gram, split).items():
        if not ms:
            continue
        scale, unit = (1e12, "TFLOP/s") if bound == "mfma" else (1e9, "GB/s")
        rate = work / (ms * 1e-3) / scale
        out[name] = {"ms": round(ms, 1), "rate": round(rate, 2), "unit": unit, "peak": peak,
                     "frac": round(rate / peak, 4), "launches": int(n)}
        weighted += ms * rate / peak
        total += ms
    return out, (round(weighted / total, 4) if total else None)


def device_ms(tm):
    return {k: round(tm[k], 1) for k in ("gemm_energy_ms", "gemm_grad_ms", "softmax_ms", "update_ms", "dots_ms",
                                         "score_ms")}


def train_network(q, seed=0):
    """The warm-start network, trained on cold-start solutions of the C1 and C2 workloads
    (prepared with the same state envelope q as the benchmarked alignment)."""
    import superdca_amd as sd
    from superdca_amd import synth, warmstart

    feats, tgts, masks = [], [], []
    for wl in ("C1", "C2"):
        st, _ = sd.prepare_alignment(synth.workload(wl)["chars"], q=q)
        f, t, m, _ = warmstart.training_set(st, q=q, n_loci=400, top=32, rand=32, seed=seed)
        feats.append(f)
        tgts.append(t)
        masks.append(m)
    params, losses = warmstart.train(np.concatenate(feats), np.concatenate(tgts), np.concatenate(masks), st, q=q,
                                     epochs=40, batch=2048, lr=2e-3, seed=seed)
    return params, losses, sum(f.shape[0] for f in feats)


WARM_MODES = {
    "direction": lambda sd: sd.WARM_DIRECTION | sd.WARM_NO_FIELDS,
    "direction-fields": lambda sd: sd.WARM_DIRECTION,
    "point": lambda sd: sd.WARM_POINT,
}


def cpu_info() -> dict:
    """Host CPU model and the cores this process may use (the GPU box grants a share of them)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "cpu_count": os.cpu_count(), "usable": usable}


def fixture_error(rows: np.ndarray, fx) -> dict:
    """Score error of solved rows against the FP64 oracle's optimum on the fixture's entries."""
    got = np.take_along_axis(rows, fx["top_idx"], axis=1)
    rel = np.abs(got - fx["top_val"]) / fx["top_val"]
    mine = np.argsort(-rows, axis=1, kind="stable")[:, :200]
    overlap = float(np.mean([len(set(a.tolist()) & set(b.tolist())) / 200 for a, b in zip(mine, fx["top_idx"])]))
    return {"loci": int(fx["loci"].size), "rel_err_max": float(rel.max()), "rel_err_median": float(np.median(rel)),
            "rel_err_p99": float(np.quantile(rel, 0.99)), "top200_overlap": round(overlap, 5)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="C4", choices=["C1", "C2", "C4"])
    ap.add_argument("--q", type=int, default=0,
                    help="state envelope (softmax width): default 5 ({A,C,G,T,-}) at C4, 4 (SuperDCA's) otherwise")
    ap.add_argument("--loci-per-step", type=int, default=0, help="default: 16384 at C4, 1024 otherwise")
    ap.add_argument("--slots", type=int, default=0, help="headline loci in flight (default 3072 at C4, 512 otherwise)")
    ap.add_argument("--history", type=int, default=0,
                    help="L-BFGS history (0: the formulation's default, 32 Gram and 10 primal)")
    ap.add_argument("--groups", type=int, default=0,
                    help="headline slot groups: 1, or 2 pipelined on two streams (0: the library's default, 2)")
    ap.add_argument("--j-threshold", type=float, default=3e-5,
                    help="headline coupling threshold on ||g_J||_2 (0: ||g||_2 < 1e-3 alone)")
    ap.add_argument("--subspace-rank", type=int, default=128,
                    help="headline: subspace warm start of this rank (128; 0: cold start)")
    ap.add_argument("--subspace-threshold", type=float, default=3e-2,
                    help="the restricted solves stop at ||g||_2 below this")
    ap.add_argument("--no-cold", action="store_true", help="skip the cold-start Gram leg on the headline's loci")
    ap.add_argument("--rule-loci", type=int, default=4096,
                    help="loci of the Gram leg at ||g||_2 < 1e-3 alone (0: skip)")
    ap.add_argument("--single-group-loci", type=int, default=2048,
                    help="Gram loci solved with one slot group after the timed legs, for the kernels' rates without "
                         "the other group sharing the GPU (0: skip)")
    ap.add_argument("--primal-loci", type=int, default=1024, help="loci of the primal split-BF16 legs (0: skip)")
    ap.add_argument("--primal-slots", type=int, default=256)
    ap.add_argument("--primal-j-threshold", type=float, default=3e-6,
                    help="primal legs' coupling threshold on ||g_J||_inf")
    ap.add_argument("--partners", type=int, default=2048, help="warm-started partners per locus (primal)")
    ap.add_argument("--warm-mode", default="direction", choices=list(WARM_MODES),
                    help="direction: network blocks as the first search direction from zero; direction-fields: "
                         "the same from the J=0 field optimum; point: start at fields + blocks")
    ap.add_argument("--no-warm", action="store_true", help="skip the warm-started primal leg")
    ap.add_argument("--fp32-loci", type=int, default=512, help="loci of the primal FP32 cold leg (0: skip)")
    ap.add_argument("--no-fixture", action="store_true",
                    help="skip the fixture solves (score errors); with --warmup 0 every kernel launch of the headline "
                         "context lies in its timed region, so a profiler's per-kernel averages match the bench line")
    ap.add_argument("--cpu-loci", type=int, default=0, help="CPU baseline sample (default: 2 loci per thread)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="default: the usable host cores, at most 16")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: there is no multi-GPU path (asked for --gpus {args.gpus}). Loci are "
              "independent given X, the weights and lambda, so the locus queue would shard across GPUs, with one "
              "row exchange for the symmetrised scores.", file=sys.stderr)
        return 2

    import superdca_amd as sd
    from superdca_amd import synth

    wl = args.workload
    q = args.q or (5 if wl == "C4" else 4)
    per_step = args.loci_per_step or (16384 if wl == "C4" else 1024)
    slots = args.slots or (3072 if wl == "C4" else 512)
    jthr = args.j_threshold

    t = time.time()
    states, translation = sd.prepare_alignment(synth.workload(wl)["chars"], q=q)
    B, L = states.shape
    partners = min(args.partners, L - 1)
    log(f"[bench] workload {wl}: B={B} L={L} generated in {time.time() - t:.1f}s")
    order = np.random.default_rng(1234).permutation(L).astype(np.int32)
    need = (args.warmup * per_step // 4 + args.steps * per_step + args.rule_loci + args.single_group_loci +
            args.primal_loci + args.fp32_loci + 3 * args.primal_slots + 64)
    if need > L:
        order = np.resize(order, need)
    pos = 0
    fx_path = os.path.join(ROOT, "tests", "golden", "c4_rows64.npz")
    fx = np.load(fx_path) if wl == "C4" and q == 5 and os.path.exists(fx_path) and not args.no_fixture else None

    # ---- headline: the Gram formulation, cold start
    gram = dict(formulation=sd.FORM_GRAM, precision=sd.PREC_SPLIT_BF16, history=args.history, grad_threshold=1e-3,
                subspace_rank=args.subspace_rank, subspace_threshold=args.subspace_threshold)
    ctx = sd.Context(states, q=q, batch_loci=slots, pipeline_groups=args.groups, j_threshold=jthr, **gram)
    t = time.time()
    w, _ = ctx.reweight(0.9)
    reweight_s = time.time() - t
    beff, lh, lJ = ctx.set_weights(w)
    log(f"[bench] K_x={ctx.num_columns}, reweighting {reweight_s * 1e3:.1f} ms, B_eff={beff:.2f}")
    first_s = None
    if args.warmup > 0:
        t = time.time()
        ctx.solve(order[pos:pos + 8])  # the identical-locus counts and the workspace
        first_s = time.time() - t
        pos += 8
    for k in range(args.warmup):
        _, wst, _, _ = ctx.solve(order[pos:pos + per_step // 4])
        assert np.all(wst["nfeval"] > 1), "warmup loci stopped at their start point"
        pos += per_step // 4
    head_from = pos
    head_f, head_sub = [], []
    head_v, head_it, head_ev, head_s, pos = run_steps(ctx, order, pos, per_step, args.steps, "gram", ctx.solve, head_f,
                                                      head_sub)
    head_tm = ctx.timing()
    accuracy = None
    if fx is not None:
        rows_fx, st_fx, _, _ = ctx.solve(fx["loci"])
        accuracy = fixture_error(rows_fx, fx)
        accuracy["converged"] = int((st_fx["status"] == sd.ST_CONVERGED).sum())
        log(f"[bench] headline score error on the fixture: {accuracy}")
    ctx.close()

    # ---- the same loci from a cold start (zero): warm vs cold on the same GPU
    gram_cold = None
    if args.subspace_rank > 0 and not args.no_cold:
        with sd.Context(states, q=q, batch_loci=slots, pipeline_groups=args.groups, j_threshold=jthr,
                        **{**gram, "subspace_rank": 0}) as cc:
            cc.set_weights(w)
            cc.solve(order[pos:pos + 8])
            pos += 8
            cold_f = []
            cv, cit, cev, _, _ = run_steps(cc, order, head_from, per_step, args.steps, "gram-cold", cc.solve, cold_f)
            ctm = cc.timing()
        fh, fc = np.concatenate(head_f), np.concatenate(cold_f)
        frel = np.abs(fh - fc) / np.abs(fc)
        gram_cold = {"loci_s": round(cv, 3), "evals_median": float(np.median(cev)), "iter_median": float(np.median(cit)),
                     "rounds": ctm["n_evals"], "device_ms": device_ms(ctm),
                     "f_warm_vs_cold_rel": {"median": float(np.median(frel)), "max": float(frel.max())}}

    # ---- the same with one slot group: the kernels' rates with the GPU to themselves
    single = None
    if args.single_group_loci > 0:
        with sd.Context(states, q=q, batch_loci=min(slots // 2, args.single_group_loci), pipeline_groups=1,
                        j_threshold=jthr, **gram) as c1:
            c1.set_weights(w)
            c1.solve(order[pos:pos + 8])
            pos += 8
            v1, _, _, _, pos = run_steps(c1, order, pos, args.single_group_loci, 1, "gram-one-group", c1.solve)
            tm1 = c1.timing()
        ph1, ph1w = phase_fractions(tm1, True, True)
        single = {"loci_s": round(v1, 3), "slots": min(slots // 2, args.single_group_loci), "phases": ph1,
                  "time_weighted_frac": ph1w, "device_ms": device_ms(tm1)}

    # ---- the Gram formulation at ||g||_2 < 1e-3 alone
    gram_rule = None
    if args.rule_loci > 0:
        with sd.Context(states, q=q, batch_loci=min(slots, args.rule_loci), pipeline_groups=args.groups,
                        j_threshold=0.0, **gram) as cr:
            cr.set_weights(w)
            cr.solve(order[pos:pos + 8])
            pos += 8
            rv, rit, rev, _, pos = run_steps(cr, order, pos, args.rule_loci, 1, "gram-rule", cr.solve)
            gram_rule = {"loci_s": round(rv, 3), "evals_median": float(np.median(rev)),
                         "convergence": "||g||_2<1e-3"}
            if fx is not None:
                rows_fx, _, _, _ = cr.solve(fx["loci"])
                gram_rule["score_error"] = fixture_error(rows_fx, fx)

    # ---- the primal formulation: split-BF16, cold and warm-started on the same loci
    primal = None
    pj = args.primal_j_threshold
    prim = dict(formulation=sd.FORM_PRIMAL, precision=sd.PREC_SPLIT_BF16, history=args.history,
                grad_threshold=1e-3, j_threshold=pj, fine_factor=0, batch_loci=args.primal_slots)
    if args.primal_loci > 0:
        net = None
        if not args.no_warm:
            t = time.time()
            net = train_network(q)
            log(f"[bench] network trained on {net[2]} pairs in {time.time() - t:.1f}s, loss {net[1][0]:.3e} -> "
                f"{net[1][-1]:.3e}")
        with sd.Context(states, q=q, **prim) as cp:
            cp.set_weights(w)
            cp.solve(order[pos:pos + args.primal_slots])  # allocation, autotuning, first launches
            pos += args.primal_slots
            from_ = pos
            cold_f = []
            pc_v, pc_it, pc_ev, _, pos = run_steps(cp, order, pos, args.primal_loci, 1, "primal-cold", cp.solve, cold_f)
            pc_tm = cp.timing()
            primal = {"convergence": f"||g||_inf<1e-3 and ||g_J||_inf<{pj:g}", "slots": args.primal_slots,
                      "cold_split_loci_s": round(pc_v, 3), "iter_median_cold": float(np.median(pc_it)),
                      "evals_median_cold": float(np.median(pc_ev)), "device_ms_cold": device_ms(pc_tm),
                      "gemm_config": cp.gemm_config()}
            if fx is not None:
                rows_fx, _, _, _ = cp.solve(fx["loci"])
                primal["score_error_cold"] = fixture_error(rows_fx, fx)
            if net is not None:
                cp.mlp_set_params(net[0], 128)
                t = time.time()
                cp.mi_topk(partners)
                mi_s = time.time() - t
                mode = WARM_MODES[args.warm_mode](sd)
                solve_w = lambda loci: cp.solve_mlp(loci, partners, mode)
                solve_w(order[pos:pos + args.primal_slots // 4])
                pos += args.primal_slots // 4
                warm_f = []
                pw_v, pw_it, pw_ev, _, _ = run_steps(cp, order, from_, args.primal_loci, 1, "primal-warm", solve_w,
                                                     warm_f)
                fc, fw = np.concatenate(cold_f), np.concatenate(warm_f)
                frel = np.abs(fw - fc) / np.abs(fc)
                primal.update({"warm_split_loci_s": round(pw_v, 3), "warm_over_cold_split": round(pw_v / pc_v, 3),
                               "iter_median_warm": float(np.median(pw_it)),
                               "evals_median_warm": float(np.median(pw_ev)),
                               "f_warm_vs_cold_rel": {"median": float(np.median(frel)), "max": float(frel.max())},
                               "partners": partners, "warm_mode": args.warm_mode, "mi_table_s": round(mi_s, 2)})

    fp32 = None
    if args.fp32_loci > 0:
        with sd.Context(states, q=q, **{**prim, "precision": sd.PREC_FP32}) as c32:
            c32.set_weights(w)
            c32.solve(order[pos:pos + args.primal_slots // 4])
            pos += args.primal_slots // 4
            fv, fit, _, _, pos = run_steps(c32, order, pos, args.fp32_loci, 1, "primal-fp32", c32.solve)
            fp32 = {"loci_s": round(fv, 3), "iter_median": float(np.median(fit)), "device_ms": device_ms(c32.timing())}

    rl = roofline(head_tm, True, True)
    phases, phases_weighted = phase_fractions(head_tm, True, True)

    cpu = None
    if not args.no_cpu:
        import oracle

        info = cpu_info()
        nthreads = args.cpu_threads or min(16, info["usable"])
        ncpu = args.cpu_loci or 2 * nthreads
        sample = order[:ncpu]
        t = time.perf_counter()
        _, _, cst = oracle.solve_loci(states, w, lh, lJ, sample, q=q, gtol=1e-3, max_iter=2000, m=10,
                                      nthreads=nthreads, blocked=True, want_rows=True)
        cpu_s = time.perf_counter() - t
        t = time.perf_counter()
        _, _, cst1 = oracle.solve_loci(states, w, lh, lJ, sample[:1], q=q, gtol=1e-3, max_iter=2000, m=10,
                                       nthreads=1, blocked=True, want_rows=True)
        cpu1_s = time.perf_counter() - t
        cpu = {"value": round(int((cst["status"] == 0).sum()) / cpu_s, 4), "unit": "loci/s", "cores": nthreads,
               "kind": "port",
               "sample": f"{ncpu} loci of the same workload at full B={B}, L={L}, q={q}, FP64 block-compressed oracle, "
                         f"cold start, SuperDCA's rule ||g||_inf<1e-3, one locus per thread on {nthreads} threads, "
                         f"{cpu_s:.1f}s wall",
               "one_thread_loci_s": round(int((cst1["status"] == 0).sum()) / cpu1_s, 4),
               "cpu_model": info["model"], "host_cpu_count": info["cpu_count"], "usable_cores": info["usable"],
               "evals_median": float(np.median(cst["nfeval"]))}
        log(f"[bench] cpu baseline {cpu}")

    conv = f"||g||_2<1e-3 and ||g_J||_2<{jthr:g}" if jthr > 0 else "||g||_2<1e-3"
    detail = {
        "formulation": "gram",
        "start": (f"subspace warm start (rank {args.subspace_rank}, restricted solves to ||g||_2 < "
                  f"{args.subspace_threshold:g})") if args.subspace_rank > 0 else "cold",
        "convergence": conv,
        "gram_cold": gram_cold,
        "warm_over_cold_gram": round(head_v / gram_cold["loci_s"], 3) if gram_cold else None,
        "evals_sub_median_headline": float(np.median(np.concatenate(head_sub))) if head_sub else None,
        "subspace_ms_headline": round(head_tm["subspace_ms"], 1),
        "score_error_headline": accuracy,
        "first_call_s": round(first_s, 3) if first_s is not None else None,
        "evals_median_headline": float(np.median(head_ev)),
        "iter_median_headline": float(np.median(head_it)),
        "iter_p95_headline": float(np.percentile(head_it, 95)),
        "rounds_headline": head_tm["n_evals"],
        "device_ms_headline": device_ms(head_tm),
        "roofline_phases_headline": phases,
        "roofline_time_weighted_frac": phases_weighted,
        "single_group": single,
        "gram_rule": gram_rule,
        "primal_split": primal,
        "cold_fp32_loci_s": fp32["loci_s"] if fp32 else None,
        "cold_fp32": fp32,
        "headline_over_cold_fp32": round(head_v / fp32["loci_s"], 3) if fp32 else None,
        "headline_over_primal_cold_split": round(head_v / primal["cold_split_loci_s"], 3) if primal else None,
        "reweight_ms": round(reweight_s * 1e3, 2), "B_eff": beff,
    }
    start_desc = (f"warm start from the locus' solve in the dominant {args.subspace_rank}-dim eigen-subspace of the "
                  "identical-locus counts") if args.subspace_rank > 0 else "cold start"
    out = {
        "metric": METRIC,
        "value": round(head_v, 4),
        "unit": "loci/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(per_step / head_v * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{wl}: B={B} sequences x N={L} loci, q_env={q}, {start_desc}, Gram formulation "
                               f"(FP64), {per_step} loci per step ({slots} in flight), L-BFGS m={args.history or 32}, {conv}",
                   "B": B, "N": L, "q_env": q, "loci_per_step": per_step, "slots": slots, "formulation": "gram",
                   "precision": "fp64", "start": start_desc, "convergence": conv,
                   "parallelism": "loci batched on 1 GPU"},
        "roofline": rl,
        "cpu_baseline": cpu,
        "detail": detail,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
