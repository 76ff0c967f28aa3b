#!/usr/bin/env python3
"""Benchmark: MFMA-filtered epsilon self-join on one MI355X.

Workload (the configuration the headline metric is quoted on): self-join of
n = 1,000,000 uniform [0,1) FP64 points in d = 64 at eps = 0.2 (seed 4),
counting unordered pairs with ||p - q|| <= eps.  One step = one complete join
(grid keys, radix sort, work list, distance kernel, band recheck) of the
points already resident in HBM.  The same join on the plain FP64 path is
timed beside it, and the CPU restatement of the reference algorithm is timed
on a bounded sample on the host cores.

Prints one JSON line.  --gpus N with N > 1 exits: this build targets one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "join s & dist-evals/s, n=1M d=64; MFMA-filter vs FP64 path; serving p99"
UNIT = "pair-distance evaluations/s"
PEAK = {"f16": 2.5e15, "bf16": 2.5e15, "fp64": 78.6e12}
PEAK_UNIT = "TFLOP/s"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--d", type=int, default=64)
    ap.add_argument("--eps", type=float, default=0.2)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--kernel", default="f16", choices=["f16", "bf16", "fp64"])
    ap.add_argument("--active-dims", type=int, default=-1, help="-1: chosen from the data")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-path comparison")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target duration of the CPU baseline sample")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run the library's streaming copy kernel over BYTES (HBM byte "
                         "counter calibration) and exit")
    return ap.parse_args()


def calibrate_copy(nbytes: int) -> None:
    import numpy as np

    import hilbertjoin_amd as hj
    rows = nbytes // (8 * 64)
    ctx = hj.Context(0)
    src = ctx.to_device(np.zeros((rows, 64)))
    dst = hj.DeviceArray(ctx, (rows, 64))
    t0 = time.perf_counter()
    ctx.device_copy(dst, src)
    dt = time.perf_counter() - t0
    print(json.dumps({"copy_bytes": src.nbytes, "read_bytes": src.nbytes,
                      "write_bytes": src.nbytes, "wall_s": dt}))


def time_join(ctx, dev, eps, kernel, steps, warmup, active):
    import hilbertjoin_amd as hj  # noqa: F401
    for _ in range(warmup):
        ctx.self_join_count(dev, eps, kernel=kernel, active_dims=active)
    walls, stats, count = [], [], None
    t_all = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        c, st = ctx.self_join_count(dev, eps, kernel=kernel, active_dims=active)
        walls.append(time.perf_counter() - t0)
        stats.append(st)
        if count is None:
            count = c
        elif c != count:
            raise RuntimeError(f"count changed between steps: {count} vs {c}")
    total = time.perf_counter() - t_all
    return count, walls, stats, total


def roofline(stats, d, kernel):
    # algorithmic work of one launch: 2*d flops per candidate pair (MFMA
    # filter) or 2*d_eff per candidate pair with d_eff the dimensions the FP64
    # kernel actually evaluated (pair_dims / candidate pairs)
    k_ms = sum(s.kernel_ms for s in stats) / len(stats)
    cand = stats[-1].candidate_pairs
    flops = 2.0 * d * cand if kernel != "fp64" else 2.0 * stats[-1].pair_dims
    achieved = flops / (k_ms * 1e-3) / 1e12
    peak = PEAK[kernel] / 1e12
    return {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak,
            "unit": PEAK_UNIT, "frac": round(achieved / peak, 5), "traffic": None,
            "kernel_ms": round(k_ms, 4), "flops_per_launch": flops,
            "flops_per_unit": f"2*d = {2 * d} per candidate pair" if kernel != "fp64" else
            f"2*d_eff per candidate pair, d_eff = {stats[-1].pair_dims / max(cand, 1):.2f}"}


def cpu_baseline(X, eps, target_s):
    """The oracle's CPU EGO join (restatement of the reference algorithm):
    sort and index over all n points, the join loop over every stride-th
    8-point block of the sorted set; the rate is extrapolated to the whole
    join."""
    import oracle
    n = X.shape[0]
    threads = oracle.max_threads()
    # calibrate the stride on a short sample
    stride = 4096
    c, t, ev = oracle.ego_self_count(X, eps, active_dims=3, stride=stride)
    if t[2] > 0:
        want = max(target_s - t[0] - t[1], 1.0)
        stride = max(1, int(stride * t[2] / want))
    c, t, ev = oracle.ego_self_count(X, eps, active_dims=3, stride=stride)
    join_full = t[2] * stride
    total = t[0] + t[1] + join_full
    pairs = n * (n - 1) / 2
    return {"value": pairs / total, "unit": UNIT, "cores": threads, "kind": "port",
            "sample": (f"CPU EGO join of the same {n} points (sort {t[0]:.2f} s + index "
                       f"{t[1]:.2f} s in full; join loop on 1/{stride} of the 8-point blocks, "
                       f"{t[2]:.2f} s, {ev} pair evaluations, extrapolated x{stride})"),
            "join_s_extrapolated": total, "sampled_evaluations": ev,
            "evaluations_per_s": ev / t[2] if t[2] > 0 else None}


def main():
    a = parse()
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus must be 1",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import hilbertjoin_amd as hj
    X = hj.gen_uniform(a.seed, a.n, a.d)
    ctx = hj.Context(0)
    cus, arch = ctx.device_info()
    dev = ctx.to_device(X)  # inputs resident in HBM before the timed region
    pairs_total = a.n * (a.n - 1) / 2

    count, walls, stats, total = time_join(ctx, dev, a.eps, a.kernel, a.steps, a.warmup,
                                           a.active_dims)
    ms = total / a.steps * 1e3
    st = stats[-1]
    out = {
        "metric": METRIC, "value": pairs_total / (ms * 1e-3), "unit": UNIT,
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": a.kernel if a.kernel != "fp64" else "f64",
        "data": f"synthetic: uniform [0,1) FP64, seed {a.seed} (counter-based SplitMix64)",
        "config": {"workload": f"self-join n={a.n} d={a.d} eps={a.eps} uniform, "
                               f"{a.kernel} distance kernel", "n": a.n, "d": a.d,
                   "eps": a.eps, "kernel": a.kernel, "active_dims": st.active_dims,
                   "key_dims": st.key_dims, "tile": 64},
        "count": count,
        "join_s": round(ms * 1e-3, 6),
        "executed_pairs_per_s": st.candidate_pairs / (ms * 1e-3),
        "candidate_pairs": st.candidate_pairs, "load_fraction": round(st.load_fraction, 6),
        "tile_pairs": st.tile_pairs, "band_pairs": st.band_pairs,
        "phase_ms": {"grid": round(st.grid_ms, 4), "sort": round(st.sort_ms, 4),
                     "index": round(st.index_ms, 4), "join": round(st.join_ms, 4),
                     "kernel": round(st.kernel_ms, 4)},
        "roofline": roofline(stats, a.d, a.kernel),
        "device": {"name": arch, "cus": cus},
    }
    if not a.no_fp64 and a.kernel != "fp64":
        c64, w64, s64, t64 = time_join(ctx, dev, a.eps, "fp64", a.steps, a.warmup,
                                       a.active_dims)
        if c64 != count:
            raise RuntimeError(f"FP64 path count {c64} != filter count {count}")
        ms64 = t64 / a.steps * 1e3
        out["fp64_path"] = {"ms_per_step": round(ms64, 4), "value": pairs_total / (ms64 * 1e-3),
                            "kernel_ms": round(sum(s.kernel_ms for s in s64) / len(s64), 4),
                            "pair_dims_per_pair": s64[-1].pair_dims /
                            max(s64[-1].candidate_pairs, 1),
                            "roofline": roofline(s64, a.d, "fp64")}
        out["speedup_vs_fp64"] = round(ms64 / ms, 3)
    out["serving_p99_ms"] = None
    dev.free()
    ctx.close()
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(X, a.eps, a.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
