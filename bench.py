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
PROFILED_TRAFFIC = {("tile", 4, "fp32", 105, 103, 1024): 152.3e6}


def build_batch(sa):
    lps = []
    for si, (m, n) in enumerate(SHAPES):
        lps += sa.gen_config(2, SEED, m, n, PER_SHAPE, shape_idx=si, batch_count=PER_SHAPE)
    return lps


def autotune(sa, precision, lps, reps=2):
    """Measured routing (blpg_autotune) on the workload's own LPs; the table is imported by
    the timed solvers.  Untimed, like any calibration that a deployment caches."""
    S = sa.Solver(precision=precision)
    table = S.autotune(lps, reps=reps)
    text = S.tune_export()
    S.close()
    names = {1: "stream", 2: "lds", 3: "tile"}
    return text, [[m, n, names[k], q, round(ms, 4)] for (_, m, n, k, q, ms) in table]


def timed_run(sa, precision, lps, steps, warmup, verify=True, model=None, tune=None):
    S = model.solver if model is not None else sa.Solver(precision=precision, verify=verify)
    if model is not None:
        S.set(precision=precision, verify=int(verify))
    if tune:
        S.tune_import(tune)
    batch = S.upload(lps)
    if model is not None:
        batch.set_warmstart(model)
    for _ in range(warmup):
        batch.solve()  # the first solve captures the batch's hipGraph; later ones replay it
        batch.fetch_raw()
    batch.sync()
    t0 = time.perf_counter()
    totals = []
    for _ in range(steps):
        batch.solve()
        batch.fetch_raw()  # waits for the solve; results land in pinned host memory
        totals.append(batch.info()["total_ms"])
    batch.sync()
    t1 = time.perf_counter()
    out = batch.fetch()
    # per-stage timings (untimed for the throughput): eager solves with timing events
    S.set(use_graph=0)
    infos = []
    for _ in range(max(2, min(steps, 5))):
        batch.solve()
        batch.sync()
        infos.append(batch.info())
    S.set(use_graph=1)
    nb = len(infos[0]["bins"])
    bins_ms = [float(np.mean([inf["bins"][i]["kernel_ms"] for inf in infos])) for i in range(nb)]
    meta = {"bins": infos[0]["bins"], "bins_ms": bins_ms,
            "verify_ms": float(np.mean([i["verify_ms"] for i in infos])),
            "fallback_ms": float(np.mean([i["fallback_ms"] for i in infos])),
            "gpu_total_ms": float(np.mean(totals)),
            "gpu_total_ms_eager": float(np.mean([i["total_ms"] for i in infos])),
            "warmstart_ms": float(np.mean([i["warmstart_ms"] for i in infos])),
            "fallback_count": int(out.fallback.sum())}
    batch.free()
    if model is None:
        S.close()
    return out, (t1 - t0) / steps, meta


def warmstart_leg(sa, steps, warmup, do_tune=True):
    """Parametric families: 64 families, A_f 105x103 fixed, 1024 instances each (b, c scaled
    by 1 + 0.05 N(0,1), seed 4).  The network is trained on the GPU on instances 0-511
    (bases from the cold FP64 solver); instances 512-1023 (32768 LPs) are timed warm-started
    in FP32 (inference included) against the cold FP64 solver on the same LPs."""
    nf, m, n = 64, 105, 103
    train, test = [], []
    for f in range(nf):
        train += sa.gen_config(4, 4, m, n, 512, shape_idx=f, first=0, batch_count=1024)
        test += sa.gen_config(4, 4, m, n, 512, shape_idx=f, first=512, batch_count=1024)
    S = sa.Solver(precision="fp64")
    t0 = time.perf_counter()
    rt = S.solve(train)
    t_label = time.perf_counter() - t0
    ws = sa.WarmStart(S, nf, m, n, seed=4, steps=1000, lr=3e-3, weight_decay=1e-2)
    t0 = time.perf_counter()
    loss = ws.train(train, rt.basis)
    t_train = time.perf_counter() - t0
    t64 = t32 = None
    tab = {}
    if do_tune:
        t64, tab["fp64"] = autotune(sa, "fp64", test[:4096])
        t32, tab["fp32"] = autotune(sa, "fp32", test[:4096])
    cold, t_cold, mc = timed_run(sa, "fp64", test, steps, warmup, verify=False, tune=t64)
    coldv, t_coldv, mcv = timed_run(sa, "fp64", test, steps, warmup, verify=True, tune=t64)
    warm, t_warm, mw = timed_run(sa, "fp32", test, steps, warmup, verify=True, model=ws, tune=t32)
    opt = cold.status == sa.OPTIMAL
    rel = np.abs(warm.objective[opt] - cold.objective[opt]) / np.maximum(1, np.abs(cold.objective[opt]))
    good = (warm.status == cold.status) & np.where(opt, np.abs(warm.objective - cold.objective)
                                                   / np.maximum(1, np.abs(cold.objective)) <= 1e-6, True)
    ws.close()
    S.close()
    return {
        "workload": "64 parametric families x 512 unseen instances (105x103), seed 4",
        "warm_fp32_lps_per_s": int(good.sum()) / t_warm, "warm_ms_per_step": t_warm * 1e3,
        "cold_fp64_lps_per_s": len(test) / t_cold, "cold_ms_per_step": t_cold * 1e3,
        "cold_fp64_checked_lps_per_s": len(test) / t_coldv,
        "speedup_vs_cold_fp64": (int(good.sum()) / t_warm) / (len(test) / t_cold),
        "speedup_vs_cold_fp64_checked": (int(good.sum()) / t_warm) / (len(test) / t_coldv),
        "accurate": int(good.sum()), "max_rel_obj_err": float(rel.max()) if opt.any() else 0.0,
        "pivots_warm": float(warm.pivots.mean()), "pivots_cold": float(cold.pivots.mean()),
        "fallbacks": mw["fallback_count"], "inference_ms": mw["warmstart_ms"],
        "warm_kernel_ms": mw["bins_ms"], "warm_verify_ms": mw["verify_ms"],
        "cold_kernel_ms": mc["bins_ms"], "train_s": t_train, "label_solve_s": t_label,
        "loss_first": float(loss[0]), "loss_last": float(loss[-1]), "tuning": tab,
    }


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


def cpu_baseline(lps, threads, min_wall=0.75):
    """The oracle over the whole batch, repeated until about min_wall seconds of wall time
    (10-20 s of CPU time on 16 threads).  Returns the last result, the wall time per batch
    and the number of repetitions."""
    import oracle
    t0 = time.perf_counter()
    reps = 0
    while True:
        r = oracle.solve_batch(lps, threads=threads)
        reps += 1
        if time.perf_counter() - t0 >= min_wall:
            break
    dt = (time.perf_counter() - t0) / reps
    return r, dt, reps


def serving_leg(sa, tune32=None, seconds=2.0, rate=50000.0):
    """Serving loop: Poisson arrivals at `rate` LPs/s of single LPs, shapes uniform over
    27x32, 50x48, 74x83, 105x103, 205x203 (generators as the first two workloads, seed 5),
    FP32 tableau + FP64 check; latency counted from each scheduled arrival.  Two seconds of
    arrivals: the drain after the last arrival (tens of ms, the largest shape's latency) is
    a fixed cost that a short window would count as lost throughput."""
    shapes = [(27, 32), (50, 48), (74, 83), (105, 103), (205, 203)]
    pools = [sa.gen_config(5, 5, m, n, 512, shape_idx=i, batch_count=1024)
             for i, (m, n) in enumerate(shapes)]
    rng = np.random.default_rng(5)
    count = int(rate * seconds)
    which = rng.integers(0, len(shapes), count)
    slot = rng.integers(0, 512, count)
    lps = [pools[w][k] for w, k in zip(which, slot)]
    arrival = np.cumsum(rng.exponential(1e6 / rate, count))
    S = sa.Solver(precision="fp32")
    if tune32:
        S.tune_import(tune32)
    srv = sa.Server(S, max_batch=2048, max_wait_us=200, precision="fp32")
    srv.replay([p[0] for p in pools] * 2, np.arange(10) * 10.0)  # arenas + graphs built
    base = srv.stats()
    t0 = time.perf_counter()
    res = srv.replay(lps, arrival)
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    S.close()
    ref = sa.Solver(precision="fp64").solve(lps[:4096])
    opt = ref.status == sa.OPTIMAL
    rel = (np.abs(res["objective"][:4096][opt] - ref.objective[opt]) /
           np.maximum(1, np.abs(ref.objective[opt])))
    lat = res["latency_us"]
    return {"workload": f"Poisson {rate:.0f} LPs/s for {seconds} s, shapes uniform over "
                        "27x32, 50x48, 74x83, 105x103, 205x203, seed 5",
            "offered": count, "completed": int((res["status"] != sa.UNDECIDED).sum()),
            # sustained rate: LPs over the span from the replay's start to the last completion
            # (arrival + latency; the Python marshalling of the requests is outside it)
            "lps_per_s": count / (float(np.max(arrival + lat)) * 1e-6),
            "span_s": float(np.max(arrival + lat)) * 1e-6, "wall_s": wall,
            "p50_us": float(np.percentile(lat, 50)), "p99_us": float(np.percentile(lat, 99)),
            "max_us": float(lat.max()), "batches": int(st["batches"] - base["batches"]),
            "mean_batch": (count / max(1, st["batches"] - base["batches"])),
            "status_agreement_fp64": float(np.mean(res["status"][:4096] == ref.status)),
            "max_rel_obj_err": float(rel.max()) if opt.any() else 0.0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="static routing (no autotune pass)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop leg")
    ap.add_argument("--no-warmstart", action="store_true",
                    help="skip the learned warm start leg (parametric families)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: LPs are independent, so the batch shards "
              "across GPUs with no exchange step; run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import simplexprojects_amd as sa

    lps = build_batch(sa)
    count = len(lps)
    tune64 = tune32 = None
    tuning = {}
    if not args.no_tune:
        tune64, tuning["fp64"] = autotune(sa, "fp64", lps)
        tune32, tuning["fp32"] = autotune(sa, "fp32", lps)
    r64, t64, meta64 = timed_run(sa, "fp64", lps, args.steps, args.warmup, tune=tune64)
    r32, t32, meta32 = timed_run(sa, "fp32", lps, args.steps, args.warmup, tune=tune32)

    # accuracy of the FP32 path against the all-FP64 run
    same = r32.status == r64.status
    opt = r64.status == sa.OPTIMAL
    rel = np.zeros(count)
    rel[opt] = np.abs(r32.objective[opt] - r64.objective[opt]) / np.maximum(1.0, np.abs(r64.objective[opt]))
    good = same & (~opt | (rel <= 1e-6))
    decided64 = np.isin(r64.status, [sa.OPTIMAL, sa.INFEASIBLE, sa.UNBOUNDED])
    n_good = int(np.sum(good & decided64))
    value = n_good / t32

    # roofline of the dominant solve kernel of the FP32 run: the bin with the most
    # algorithmic bytes (its duration comes from the HIP events of the eager solves, where
    # the bins share the GPU as they do in the timed graph replays)
    def bin_lps(b):
        shape = (meta32["bins"][b]["rcap"] - 2, meta32["bins"][b]["ccap"] - 2)
        return shape, [k for k, lp in enumerate(lps) if (lp.m, lp.n) == shape]
    bi = int(np.argmax([algorithmic_bytes(lps, r32, bin_lps(b)[1], 4)
                        for b in range(len(meta32["bins"]))]))
    bin_shape, idx = bin_lps(bi)
    bytes_launch = algorithmic_bytes(lps, r32, idx, 4)
    flops_launch = algorithmic_flops(lps, r32, idx)
    ms = meta32["bins_ms"][bi]
    achieved = bytes_launch / (ms * 1e-3) / 1e9
    kname = {1: "stream", 2: "lds", 3: "tile"}[meta32["bins"][bi]["kernel"]]
    kparam = meta32["bins"][bi]["param"]

    cpu = None
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        rc, dtc, reps = cpu_baseline(lps, threads)
        agree = float(np.mean(rc["status"] == r64.status))
        cpu = {"value": count / dtc, "unit": "LPs/s", "cores": threads, "kind": "port",
               "sample": f"the full {count}-LP batch {reps} times, oracle FP64 replay of the "
                         f"same pivot rule (C, -O3 AVX2/FMA, OpenMP), {dtc * reps:.2f} s wall "
                         f"({dtc * reps * threads:.0f} thread-seconds); status agreement with "
                         f"GPU FP64 {agree:.4f}"}

    wsleg = (None if args.no_warmstart else
             warmstart_leg(sa, args.steps, args.warmup, do_tune=not args.no_tune))

    serve = None if args.no_serve else serving_leg(sa, tune32)

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
                     "traffic": PROFILED_TRAFFIC.get((kname, kparam, "fp32") + bin_shape + (len(idx),)),
                     "kernel": f"{kname}[{kparam}] fp32 bin {bin_shape[0]}x{bin_shape[1]}",
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
        "tuning": tuning,
        "warmstart": wsleg,
        "serving": serve,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
