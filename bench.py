#!/usr/bin/env python3
"""bench.py -- batched LP throughput on one MI355X.

Workload (one "step"): the mixed-size batch of 4096 inequality LPs, 1024 each at
50x48, 56x97, 74x83 and 105x103 (A ~ U[-1,1) with 40% zeroed, b ~ U[0,10) with 10% of
rows exactly 0, c ~ N(0,1), 5% planted unbounded; seed 2), already resident in HBM,
solved through the C ABI (blpg_batch_solve: size-bin kernels, FP64 re-solve + KKT check,
FP64 fallback) with the results copied back to the host (blpg_batch_download).

`value` counts the LPs of the FP32-tableau run whose status equals the all-FP64 run's and,
when optimal, whose objective is within 1e-6 * max(1, |obj_fp64|); the all-FP64 run is
timed the same way and reported beside it.  The roofline object covers the dominant
solve kernel: algorithmic bytes per launch (sum over its LPs of pivots * 2*(m+1)*(n+1)*s,
one more row per phase-I pivot, plus the LP's input and output bytes) over its average
HIP-event duration.  The CPU baseline is the oracle's FP64 replay of the same rule over
the host's cores on the same LPs.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "batched LPs/sec at ≤1e-6 rel obj err (warm FP32 vs plain FP64); HBM GB/s frac"
SHAPES = [(50, 48), (56, 97), (74, 83), (105, 103)]
PER_SHAPE = 1024
SEED = 2
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by the
# same counters on a plain 256 MiB device copy (FETCH x2.000, WRITE x1.000); the commands and
# raw counts are in profiles/tile_fp32_105x103_mixed_batch.md.
PROFILED_TRAFFIC = {("tile", "fp32", 105, 103, 1024): 92.65e6}


def build_batch(sa):
    lps = []
    for si, (m, n) in enumerate(SHAPES):
        lps += sa.gen_config(2, SEED, m, n, PER_SHAPE, shape_idx=si, batch_count=PER_SHAPE)
    return lps


def timed_run(sa, precision, lps, steps, warmup):
    S = sa.Solver(precision=precision, verify=True)
    batch = S.upload(lps)
    res, xs, bs = sa._alloc_results(lps)
    lib = S.lib
    bins_ms = None
    for _ in range(warmup):
        batch.solve()
        S._check(lib.blpg_batch_download(S.ctx, batch.h, res))
    S._check(lib.blpg_batch_sync(S.ctx))
    t0 = time.perf_counter()
    infos = []
    for _ in range(steps):
        batch.solve()
        S._check(lib.blpg_batch_download(S.ctx, batch.h, res))  # includes the stream sync
        infos.append(batch.info())
    S._check(lib.blpg_batch_sync(S.ctx))
    t1 = time.perf_counter()
    out = sa._collect(res, xs, bs)
    nb = len(infos[0]["bins"])
    bins_ms = [float(np.mean([inf["bins"][i]["kernel_ms"] for inf in infos])) for i in range(nb)]
    meta = {"bins": infos[0]["bins"], "bins_ms": bins_ms,
            "verify_ms": float(np.mean([i["verify_ms"] for i in infos])),
            "fallback_ms": float(np.mean([i["fallback_ms"] for i in infos])),
            "gpu_total_ms": float(np.mean([i["total_ms"] for i in infos])),
            "fallback_count": int(out.fallback.sum())}
    batch.free()
    S.close()
    return out, (t1 - t0) / steps, meta


def algorithmic_bytes(lps, res, idx, s):
    """Sum over LPs of pivots * 2*(m+1)*(n+1)*s (+1 row in phase I) + I/O bytes."""
    tot = 0
    for k in idx:
        m, n = lps[k].m, lps[k].n
        p = int(res.pivots[k]); p1 = int(res.phase1_pivots[k])
        tot += (p - p1) * 2 * (m + 1) * (n + 1) * s + p1 * 2 * (m + 2) * (n + 1) * s
        tot += (m * n + m + n) * 8 + (n + m) * (8 + 4) + 16
    return tot


def algorithmic_flops(lps, res, idx):
    tot = 0
    for k in idx:
        m, n = lps[k].m, lps[k].n
        p = int(res.pivots[k]); p1 = int(res.phase1_pivots[k])
        tot += (p - p1) * 2 * (m + 1) * (n + 1) + p1 * 2 * (m + 2) * (n + 1)
    return tot


def cpu_baseline(lps, threads):
    import oracle
    t0 = time.perf_counter()
    r = oracle.solve_batch(lps, threads=threads)
    dt = time.perf_counter() - t0
    return r, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: LPs are independent, so the batch shards "
              "across GPUs with no exchange step; run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import simplexprojects_amd as sa

    lps = build_batch(sa)
    count = len(lps)
    r64, t64, meta64 = timed_run(sa, "fp64", lps, args.steps, args.warmup)
    r32, t32, meta32 = timed_run(sa, "fp32", lps, args.steps, args.warmup)

    # accuracy of the FP32 path against the all-FP64 run
    same = r32.status == r64.status
    opt = r64.status == sa.OPTIMAL
    rel = np.zeros(count)
    rel[opt] = np.abs(r32.objective[opt] - r64.objective[opt]) / np.maximum(1.0, np.abs(r64.objective[opt]))
    good = same & (~opt | (rel <= 1e-6))
    decided64 = np.isin(r64.status, [sa.OPTIMAL, sa.INFEASIBLE, sa.UNBOUNDED])
    n_good = int(np.sum(good & decided64))
    value = n_good / t32

    # roofline of the dominant solve kernel of the FP32 run
    bi = int(np.argmax(meta32["bins_ms"]))
    bin_shape = (meta32["bins"][bi]["rcap"] - 2, meta32["bins"][bi]["ccap"] - 2)
    idx = [k for k, lp in enumerate(lps) if (lp.m, lp.n) == bin_shape]
    bytes_launch = algorithmic_bytes(lps, r32, idx, 4)
    flops_launch = algorithmic_flops(lps, r32, idx)
    ms = meta32["bins_ms"][bi]
    achieved = bytes_launch / (ms * 1e-3) / 1e9
    kname = {1: "stream", 2: "lds", 3: "tile"}[meta32["bins"][bi]["kernel"]]

    cpu = None
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        rc, dtc = cpu_baseline(lps, threads)
        agree = float(np.mean(rc["status"] == r64.status))
        cpu = {"value": count / dtc, "unit": "LPs/s", "cores": threads, "kind": "port",
               "sample": f"the full {count}-LP batch once, oracle FP64 replay of the same "
                         f"pivot rule (C, -O3 AVX2/FMA, OpenMP), {dtc:.2f} s wall; "
                         f"status agreement with GPU FP64 {agree:.4f}"}

    line = {
        "metric": METRIC, "value": value, "unit": "LPs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t32 * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "mixed-size batch: 4096 LPs, 1024 each at 50x48, 56x97, "
                               "74x83, 105x103; seed 2; FP32 tableau + FP64 basis re-solve/"
                               "KKT check + FP64 fallback; inputs resident in HBM, results "
                               "to host",
                   "lps": count, "seed": SEED},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": PROFILED_TRAFFIC.get((kname, "fp32") + bin_shape + (len(idx),)),
                     "kernel": f"{kname} fp32 bin {bin_shape[0]}x{bin_shape[1]}",
                     "launch_ms": ms, "algorithmic_bytes": bytes_launch,
                     "fp64_valu_equiv_tflops": flops_launch / (ms * 1e-3) / 1e12},
        "cpu_baseline": cpu,
        "fp32": {"lps_per_s_all": count / t32, "accurate": n_good, "ms_per_step": t32 * 1e3,
                 "fallbacks": meta32["fallback_count"], "bins_ms": meta32["bins_ms"],
                 "verify_ms": meta32["verify_ms"], "fallback_ms": meta32["fallback_ms"],
                 "gpu_ms": meta32["gpu_total_ms"], "max_rel_obj_err": float(rel.max())},
        "fp64": {"lps_per_s": count / t64, "ms_per_step": t64 * 1e3,
                 "bins_ms": meta64["bins_ms"], "verify_ms": meta64["verify_ms"],
                 "gpu_ms": meta64["gpu_total_ms"]},
        "status_mix": {sa.STATUS_NAMES.get(int(s), str(s)): int(c)
                       for s, c in zip(*np.unique(r64.status, return_counts=True))},
        "pivots_mean": float(np.mean(r64.pivots)),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
