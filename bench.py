"""Throughput of the per-locus plmDCA solve on one MI355X: loci/s to convergence.

A step is one batch of loci (2048 at genome scale) solved from a cold start to ||g||_inf < 1e-3 (SuperDCA's default stopping rule) on the genome-scale
workload B = 3000 sequences x N = 94,000 loci (seeded synthetic alignment, 60 clades,
2 % mutation, 2000 planted pairs).  Timing starts with the alignment, weights and
lambda resident on the device and ends when the last score row is in host memory.

Prints one JSON line (rank 0): metric, value (loci/s), roofline of the dominant kernel
(measured with HIP events on the library's stream), and a CPU baseline: the FP64 oracle
(block-compressed, one locus per thread) timed on a bounded sample of the same workload.
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="C4", choices=["C1", "C2", "C4"])
    ap.add_argument("--loci-per-step", type=int, default=0, help="default: 2048 at C4, 1024 otherwise")
    ap.add_argument("--slots", type=int, default=0,
                    help="loci in flight on the GPU (default 256 at C4, 128 otherwise); finished slots are refilled")
    ap.add_argument("--precision", default="", choices=["", "fp32", "split-bf16"])
    ap.add_argument("--history", type=int, default=10)
    ap.add_argument("--cpu-loci", type=int, default=0, help="CPU baseline sample size (default: one per thread)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "see DESIGN.md for how the loci would shard.", file=sys.stderr)
        return 2

    import superdca_amd as sd
    from superdca_amd import synth

    wl = args.workload
    cfg = synth.WORKLOADS[wl]
    prec_name = args.precision or ("split-bf16" if wl == "C4" else "fp32")
    prec = sd.PREC_SPLIT_BF16 if prec_name == "split-bf16" else sd.PREC_FP32
    per_step = args.loci_per_step or (2048 if wl == "C4" else 1024)
    slots = args.slots or (256 if wl == "C4" else 128)

    t = time.time()
    data = synth.workload(wl)
    states, translation = sd.prepare_alignment(data["chars"])
    del data
    B, L = states.shape
    log(f"[bench] workload {wl}: B={B} L={L} generated in {time.time() - t:.1f}s")

    t = time.time()
    ctx = sd.Context(states, q=4, precision=prec, batch_loci=slots, history=args.history, grad_threshold=1e-3)
    log(f"[bench] context (one-hot X, K_x={ctx.num_columns}) built in {time.time() - t:.1f}s")
    t = time.time()
    w, _ = ctx.reweight(0.9)
    reweight_s = time.time() - t
    beff, lh, lJ = ctx.set_weights(w)
    log(f"[bench] reweighting {reweight_s * 1e3:.1f} ms, B_eff={beff:.2f} lambda_h={lh:.3f} lambda_J={lJ:.3f}")

    order = np.random.default_rng(1234).permutation(L).astype(np.int32)
    need = (args.warmup + args.steps) * per_step
    if need > L:
        order = np.resize(order, need)
    pos = 0
    for k in range(args.warmup):
        t = time.time()
        _, wst, _, _ = ctx.solve(order[pos:pos + per_step])
        assert np.all(wst["nfeval"] > 1), "warmup loci stopped at their start point"
        pos += per_step
        log(f"[bench] warmup step {k} {time.time() - t:.2f}s")
    ctx.timing(reset=True)
    converged = 0
    iters, evals = [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        rows, stats, _, rc = ctx.solve(order[pos:pos + per_step])
        pos += per_step
        converged += int((stats["status"] == sd.ST_CONVERGED).sum())
        iters.append(stats["iter"])
        evals.append(stats["nfeval"])
        log(f"[bench] step {k} done: {converged} converged so far, {time.perf_counter() - t0:.2f}s")
    elapsed = time.perf_counter() - t0
    tm = ctx.timing()
    iters = np.concatenate(iters)
    evals = np.concatenate(evals)
    value = converged / elapsed

    # dominant kernel group by device time
    groups = {
        "k_gemm_nt<EpiGrad>": (tm["gemm_grad_ms"], tm["flops_grad"], "mfma"),
        "k_gemm_nt<EpiEnergy>": (tm["gemm_energy_ms"], tm["flops_energy"], "mfma"),
        "k_update": (tm["update_ms"], tm["bytes_update"], "hbm"),
        "k_dots": (tm["dots_ms"], tm["bytes_dots"], "hbm"),
    }
    dom = max(groups, key=lambda k: groups[k][0])
    ms, work, bound = groups[dom]
    if bound == "mfma":
        peak = PEAK["bf16_tflops"] if prec == sd.PREC_SPLIT_BF16 else PEAK["fp32_tflops"]
        achieved = work / (ms * 1e-3) / 1e12
        unit = "TFLOP/s"
    else:
        peak = PEAK["hbm_gbs"]
        achieved = work / (ms * 1e-3) / 1e9
        unit = "GB/s"
    roofline = {"bound": bound, "kernel": dom, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
                "frac": round(achieved / peak, 5), "traffic": None}

    cpu = None
    if not args.no_cpu:
        import oracle

        nthreads = args.cpu_threads or min(16, os.cpu_count() or 1)
        ncpu = args.cpu_loci or nthreads
        sample = order[:ncpu]
        t = time.perf_counter()
        _, _, cst = oracle.solve_loci(states, w, lh, lJ, sample, gtol=1e-3, max_iter=2000, m=10, nthreads=nthreads,
                                      blocked=True, want_rows=True)
        cpu_s = time.perf_counter() - t
        cpu = {"value": round(int((cst["status"] == 0).sum()) / cpu_s, 4), "unit": "loci/s", "cores": nthreads,
               "kind": "port",
               "sample": f"{ncpu} loci of the same workload at full B={B}, L={L}, FP64 block-compressed "
                         f"oracle, one locus per thread, {cpu_s:.1f}s wall"}
        log(f"[bench] cpu baseline {cpu}")

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "loci/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x2+fp32" if prec == sd.PREC_SPLIT_BF16 else "fp32",
        "data": "synthetic",
        "config": {"workload": f"{wl}: B={B} sequences x N={L} loci, cold start, {per_step} loci per step "
                               f"({slots} in flight), L-BFGS m={args.history}, ||g||_inf<1e-3", "B": B, "N": L,
                   "loci_per_step": per_step, "slots": slots,
                   "precision": prec_name, "start": "cold", "parallelism": "loci batched on 1 GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "detail": {
            "converged": converged, "iter_median": float(np.median(iters)), "iter_p95": float(np.percentile(iters, 95)),
            "evals_median": float(np.median(evals)), "evals_p95": float(np.percentile(evals, 95)),
            "reweight_ms": round(reweight_s * 1e3, 2), "B_eff": beff,
            "device_ms": {k: round(tm[k], 2) for k in ("gemm_energy_ms", "gemm_grad_ms", "softmax_ms", "update_ms",
                                                       "dots_ms", "score_ms")},
            "kernel_rates": {k: round(v[1] / (v[0] * 1e-3) / (1e12 if v[2] == "mfma" else 1e9), 2)
                             for k, v in groups.items() if v[0] > 0},
            "rounds": tm["n_evals"],
        },
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
