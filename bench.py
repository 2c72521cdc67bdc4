"""Throughput of the per-locus plmDCA solve on one MI355X: loci/s to convergence.

Workload: B = 3000 sequences x N = 94,000 loci (seeded synthetic genome alignment, 60
clades, 2 % mutation, 2000 planted pairs), q = 5 envelope, L-BFGS m = 10.  A step is a
batch of loci solved to convergence with 256 loci in flight on the GPU (finished slots
are refilled).  Timing starts with the alignment, weights and lambda resident on the
device and ends when the last score row is in host memory.

Convergence (headline): SuperDCA's rule ||g||_inf < 1e-3 plus the coupling threshold
||g_J||_inf < 3e-6, at which the coupling scores lie within 1e-4 relative of the FP64
optimum; the bench measures that error itself on the 64 loci of the committed C4
fixture, solved with the headline settings.  SuperDCA's rule alone is reported beside it.

Legs on the same GPU (warm and cold solve the same loci, so their final objectives are
compared locus by locus):
  warm   split-BF16 contractions, warm starts from the network (trained beforehand on
         cold-start solutions of the C1 and C2 workloads, untimed) for the top-2048 MI
         partners of every locus -- the headline value
  cold   split-BF16, cold start
  fp32   the plain path: FP32 MFMA contractions, cold start, the same convergence rule
  rule   split-BF16 warm and cold at SuperDCA's rule alone
The MI partner table of the alignment is built once (reported, not in the loci/s).

Prints one JSON line: value (warm loci/s), the roofline of the dominant kernel of the
headline leg (HIP events on the library's stream), and a CPU baseline: the FP64 oracle
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

PEAK = {"bf16_tflops": 2500.0, "fp32_tflops": 157.3, "hbm_gbs": 8000.0}
METRIC = "loci/s to converged plmDCA at B=3000,N=94k; warm vs cold start, same GPU"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def run_steps(ctx, order, pos, per_step, steps, label, solve, fvals=None):
    """Times `steps` solve calls; returns (loci/s, iterations, evaluations, seconds, new pos).
    The final objective values are appended to `fvals` when given."""
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
        log(f"[bench] {label} step {k}: {converged} converged, {time.perf_counter() - t0:.2f}s")
    elapsed = time.perf_counter() - t0
    return converged / elapsed, np.concatenate(iters), np.concatenate(evals), elapsed, pos


# HBM bytes moved per algorithmic byte, from rocprofv3 FETCH_SIZE and WRITE_SIZE passes over
# the split-precision C4 solve (scaled by a plain device copy: fetch x2.000, write x1.000);
# profiles/c4_q5_solve_kernels.md holds the commands and the per-dispatch table.
HBM_PER_ALGORITHMIC_BYTE = {"k_update": 13.34 / 13.10, "k_dots": 10.51 / 10.50}


def roofline(tm, split):
    # one launch of each kernel per round (the energy and gradient launches are counted
    # separately by the library)
    groups = {
        "k_gemm_nt<EpiGrad>": (tm["gemm_grad_ms"], tm["flops_grad"], "mfma", tm["n_grad_launches"]),
        "k_gemm_nt<EpiEnergy>": (tm["gemm_energy_ms"], tm["flops_energy"], "mfma", tm["n_energy_launches"]),
        "k_update": (tm["update_ms"], tm["bytes_update"], "hbm", tm["n_evals"]),
        "k_dots": (tm["dots_ms"], tm["bytes_dots"], "hbm", tm["n_evals"]),
    }
    dom = max(groups, key=lambda k: groups[k][0])
    ms, work, bound, n = groups[dom]
    if bound == "mfma":
        peak = PEAK["bf16_tflops"] if split else PEAK["fp32_tflops"]
        achieved, unit = work / (ms * 1e-3) / 1e12, "TFLOP/s"
    else:
        peak = PEAK["hbm_gbs"]
        achieved, unit = work / (ms * 1e-3) / 1e9, "GB/s"
    rates = {k: round(v[1] / (v[0] * 1e-3) / (1e12 if v[2] == "mfma" else 1e9), 2) for k, v in groups.items() if v[0]}
    ratio = HBM_PER_ALGORITHMIC_BYTE.get(dom) if split else None
    traffic = round(ratio * work / max(n, 1)) if ratio else None
    return ({"bound": bound, "kernel": dom, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
             "frac": round(achieved / peak, 5), "traffic": traffic,
             "traffic_basis": (f"HBM bytes per launch: algorithmic bytes x {ratio:.4f}, the measured "
                               "(FETCH_SIZE + WRITE_SIZE) / algorithmic ratio of this kernel"
                               if ratio else None),
             "launches": int(n),
             "avg_launch_us": round(ms * 1e3 / max(n, 1), 1),
             "work_per_launch": round(work / max(n, 1) / (1e12 if bound == "mfma" else 1e9), 4)}, rates)


def phase_fractions(tm, split):
    """Roofline fraction of each of the four dominant kernels (algorithmic work over its
    device time, against the MFMA or HBM peak) and their device-time-weighted mean."""
    groups = [("k_gemm_nt<EpiEnergy>", "gemm_energy_ms", "flops_energy", "mfma"),
              ("k_gemm_nt<EpiGrad>", "gemm_grad_ms", "flops_grad", "mfma"),
              ("k_update", "update_ms", "bytes_update", "hbm"),
              ("k_dots", "dots_ms", "bytes_dots", "hbm")]
    out, weighted, total = {}, 0.0, 0.0
    for name, ms_key, work_key, bound in groups:
        ms, work = tm[ms_key], tm[work_key]
        if not ms:
            continue
        if bound == "mfma":
            peak, scale, unit = (PEAK["bf16_tflops"] if split else PEAK["fp32_tflops"]), 1e12, "TFLOP/s"
        else:
            peak, scale, unit = PEAK["hbm_gbs"], 1e9, "GB/s"
        rate = work / (ms * 1e-3) / scale
        out[name] = {"ms": round(ms, 1), "rate": round(rate, 2), "unit": unit, "peak": peak,
                     "frac": round(rate / peak, 4)}
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
    ap.add_argument("--loci-per-step", type=int, default=0, help="default: 2048 at C4, 1024 otherwise")
    ap.add_argument("--slots", type=int, default=0, help="loci in flight (default 256 at C4, 128 otherwise)")
    ap.add_argument("--history", type=int, default=10)
    ap.add_argument("--groups", type=int, default=1, help="slot groups pipelined on two streams (1 or 2)")
    ap.add_argument("--j-threshold", type=float, default=3e-6,
                    help="coupling-gradient threshold of the headline convergence rule (0: SuperDCA's rule alone)")
    ap.add_argument("--fine-factor", type=int, default=0,
                    help="split precision: fine level below fine_factor * j_threshold (0: none)")
    ap.add_argument("--partners", type=int, default=2048, help="warm-started partners per locus")
    ap.add_argument("--warm-mode", default="direction", choices=list(WARM_MODES),
                    help="direction: network blocks as the first search direction from zero; direction-fields: "
                         "the same from the J=0 field optimum; point: start at fields + blocks")
    ap.add_argument("--fp32-loci", type=int, default=1024, help="loci of the FP32 cold leg (0: skip)")
    ap.add_argument("--fp32-slots", type=int, default=0,
                    help="loci in flight in the FP32 leg (default: as --slots)")
    ap.add_argument("--rule-loci", type=int, default=2048,
                    help="loci of each leg at SuperDCA's rule alone (0: skip)")
    ap.add_argument("--no-warm", action="store_true", help="skip the warm legs (value = cold split-BF16)")
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
    per_step = args.loci_per_step or (2048 if wl == "C4" else 1024)
    slots = args.slots or (256 if wl == "C4" else 128)
    slots32 = args.fp32_slots or slots
    jthr = args.j_threshold

    t = time.time()
    states, translation = sd.prepare_alignment(synth.workload(wl)["chars"], q=q)
    B, L = states.shape
    partners = min(args.partners, L - 1)
    log(f"[bench] workload {wl}: B={B} L={L} generated in {time.time() - t:.1f}s")
    order = np.random.default_rng(1234).permutation(L).astype(np.int32)
    need = (2 * args.warmup + 2 * args.steps) * per_step + args.fp32_loci + 3 * args.rule_loci + 2 * max(slots, slots32)
    if need > L:
        order = np.resize(order, need)
    pos = 0
    fx_path = os.path.join(ROOT, "tests", "golden", "c4_rows64.npz")
    fx = np.load(fx_path) if wl == "C4" and q == 5 and os.path.exists(fx_path) else None

    net = None
    if not args.no_warm:
        t = time.time()
        net = train_network(q)
        log(f"[bench] network trained on {net[2]} pairs in {time.time() - t:.1f}s, loss {net[1][0]:.3e} -> "
            f"{net[1][-1]:.3e}")

    rule = dict(grad_threshold=1e-3)
    accurate = dict(grad_threshold=1e-3, j_threshold=jthr, fine_factor=args.fine_factor)
    ctx = sd.Context(states, q=q, precision=sd.PREC_SPLIT_BF16, batch_loci=slots, history=args.history,
                     pipeline_groups=args.groups, **accurate)
    t = time.time()
    w, _ = ctx.reweight(0.9)
    reweight_s = time.time() - t
    beff, lh, lJ = ctx.set_weights(w)
    log(f"[bench] K_x={ctx.num_columns}, reweighting {reweight_s * 1e3:.1f} ms, B_eff={beff:.2f}")

    # cold split-BF16
    for k in range(args.warmup):
        _, wst, _, _ = ctx.solve(order[pos:pos + per_step])
        assert np.all(wst["nfeval"] > 1), "warmup loci stopped at their start point"
        pos += per_step
    timed_from = pos  # the warm leg solves the same loci
    cold_f = []
    cold_v, cold_it, cold_ev, cold_s, pos = run_steps(ctx, order, pos, per_step, args.steps, "cold-split", ctx.solve,
                                                      cold_f)
    cold_tm = ctx.timing()
    gemm_cfg = ctx.gemm_config()
    log(f"[bench] split-BF16 GEMM configuration {gemm_cfg}")

    warm = None
    solve_head = ctx.solve
    if net is not None:
        ctx.mlp_set_params(net[0], 128)
        t = time.time()
        ctx.mi_topk(partners)
        mi_s = time.time() - t
        log(f"[bench] MI top-{partners} partner table in {mi_s:.2f}s")
        mode = WARM_MODES[args.warm_mode](sd)
        solve_w = lambda loci: ctx.solve_mlp(loci, partners, mode)
        for k in range(args.warmup):
            solve_w(order[pos:pos + per_step // 4])
            pos += per_step // 4
        warm_f = []
        warm_v, warm_it, warm_ev, warm_s, _ = run_steps(ctx, order, timed_from, per_step, args.steps, "warm-split",
                                                        solve_w, warm_f)
        fc, fw = np.concatenate(cold_f), np.concatenate(warm_f)
        frel = np.abs(fw - fc) / np.abs(fc)
        warm = {"loci_s": warm_v, "iter": warm_it, "evals": warm_ev, "tm": ctx.timing(), "mi_s": mi_s,
                "f_rel": {"median": float(np.median(frel)), "max": float(frel.max())}}
        solve_head = solve_w
    # score error of the headline settings: the fixture's loci, solved like the timed ones
    accuracy = None
    if fx is not None:
        rows_fx, st_fx, _, _ = solve_head(fx["loci"])
        accuracy = fixture_error(rows_fx, fx)
        accuracy["converged"] = int((st_fx["status"] == sd.ST_CONVERGED).sum())
        log(f"[bench] headline score error on the fixture: {accuracy}")
    ctx.close()

    # SuperDCA's rule alone (secondary figures)
    rule_legs = None
    if args.rule_loci > 0:
        ctxr = sd.Context(states, q=q, precision=sd.PREC_SPLIT_BF16, batch_loci=slots, history=args.history,
                          pipeline_groups=args.groups, **rule)
        ctxr.set_weights(w)
        ctxr.solve(order[pos:pos + slots])  # allocation and first launches
        pos += slots
        rule_from = pos
        rc_v, rc_it, _, _, pos = run_steps(ctxr, order, pos, args.rule_loci, 1, "cold-split-rule", ctxr.solve)
        rule_legs = {"cold_split_loci_s": round(rc_v, 3), "cold_split_iter_median": float(np.median(rc_it))}
        if net is not None:
            ctxr.mlp_set_params(net[0], 128)
            ctxr.mi_topk(partners)
            mode = WARM_MODES[args.warm_mode](sd)
            rw_v, rw_it, _, _, _ = run_steps(ctxr, order, rule_from, args.rule_loci, 1, "warm-split-rule",
                                             lambda loci: ctxr.solve_mlp(loci, partners, mode))
            rule_legs.update({"warm_split_loci_s": round(rw_v, 3), "warm_split_iter_median": float(np.median(rw_it)),
                              "warm_over_cold_split": round(rw_v / rc_v, 3)})
        if fx is not None:
            rows_fx, _, _, _ = ctxr.solve(fx["loci"])
            rule_legs["score_error"] = fixture_error(rows_fx, fx)
        ctxr.close()

    fp32 = None
    if args.fp32_loci > 0:
        ctx32 = sd.Context(states, q=q, precision=sd.PREC_FP32, batch_loci=slots32, history=args.history,
                           pipeline_groups=args.groups, **accurate)
        ctx32.set_weights(w)
        ctx32.solve(order[pos:pos + slots32 // 4])  # allocation and first launches
        pos += slots32 // 4
        fv, fit, fev, fs, pos = run_steps(ctx32, order, pos, args.fp32_loci, 1, "cold-fp32", ctx32.solve)
        fp32 = {"loci_s": fv, "iter": fit, "tm": ctx32.timing(), "gemm": ctx32.gemm_config()}
        ctx32.close()

    head = warm if warm is not None else {"loci_s": cold_v, "iter": cold_it, "evals": cold_ev, "tm": cold_tm}
    rl, rates = roofline(head["tm"], True)
    phases, phases_weighted = phase_fractions(head["tm"], True)

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

    detail = {
        "convergence": f"||g||_inf<1e-3 and ||g_J||_inf<{jthr:g}" if jthr > 0 else "||g||_inf<1e-3",
        "fine_factor": args.fine_factor,
        "score_error_headline": accuracy,
        "cold_split_bf16_loci_s": round(cold_v, 3),
        "cold_fp32_loci_s": round(fp32["loci_s"], 3) if fp32 else None,
        "fp32_slots": slots32 if fp32 else None,
        "warm_split_bf16_loci_s": round(warm["loci_s"], 3) if warm else None,
        "warm_over_cold_fp32": round(warm["loci_s"] / fp32["loci_s"], 3) if (warm and fp32) else None,
        "warm_over_cold_split": round(warm["loci_s"] / cold_v, 3) if warm else None,
        "superdca_rule": rule_legs,
        "iter_median": {"cold_split": float(np.median(cold_it)),
                        "warm_split": float(np.median(warm["iter"])) if warm else None,
                        "cold_fp32": float(np.median(fp32["iter"])) if fp32 else None},
        "iter_p95_headline": float(np.percentile(head["iter"], 95)),
        "evals_median_headline": float(np.median(head["evals"])),
        # same loci in both legs: final f of the warm start relative to the cold start's
        "f_warm_vs_cold_rel": warm["f_rel"] if warm else None,
        "partners": partners if warm else 0,
        "warm_mode": args.warm_mode if warm else None,
        "mi_table_s": round(warm["mi_s"], 2) if warm else None,
        "reweight_ms": round(reweight_s * 1e3, 2), "B_eff": beff,
        "device_ms_headline": device_ms(head["tm"]),
        "kernel_rates_headline": rates,
        "roofline_phases_headline": phases,
        "roofline_time_weighted_frac": phases_weighted,
        "device_ms_cold_fp32": device_ms(fp32["tm"]) if fp32 else None,
        "rounds_headline": head["tm"]["n_evals"],
        "gemm_config_split": gemm_cfg,
        "gemm_config_fp32": fp32["gemm"] if fp32 else None,
    }
    conv = f"||g||_inf<1e-3 and ||g_J||_inf<{jthr:g}" if jthr > 0 else "||g||_inf<1e-3"
    out = {
        "metric": METRIC,
        "value": round(head["loci_s"], 4),
        "unit": "loci/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(per_step / head["loci_s"] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x2+fp32",
        "data": "synthetic",
        "config": {"workload": f"{wl}: B={B} sequences x N={L} loci, q_env={q}, {'warm' if warm else 'cold'} start, "
                               f"{per_step} loci per step ({slots} in flight), L-BFGS m={args.history}, {conv}",
                   "B": B, "N": L, "q_env": q, "loci_per_step": per_step, "slots": slots,
                   "precision": "split-bf16", "start": "warm" if warm else "cold", "convergence": conv,
                   "parallelism": "loci batched on 1 GPU"},
        "roofline": rl,
        "cpu_baseline": cpu,
        "detail": detail,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
