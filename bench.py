#!/usr/bin/env python3
"""Benchmark: MFMA-filtered epsilon self-join on one MI355X.

Workload (the configuration the headline metric is quoted on): self-join of
n = 1,000,000 uniform [0,1) FP64 points in d = 64 at eps = 0.2 (seed 4),
counting unordered pairs with ||p - q|| <= eps.  One step = one complete join
(grid keys, radix sort, work list, distance kernel, band recheck) of the
points already resident in HBM.  The same join on the plain FP64 path is
timed beside it, and the CPU restatement of the reference algorithm is timed
on a bounded sample on the host cores.

Prints one JSON line.  --gpus N with N > 1 exits: the join runs on one GPU.
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
# value: the whole join's pair count n(n-1)/2 per second of join wall time
# (brute-force-equivalent pair-distance evaluations/s); the pairs the distance
# kernel actually evaluates are reported beside it (executed_pairs_per_s)
UNIT = "brute-force-equivalent pair-distance evaluations/s (n(n-1)/2 per join second)"
PEAK = {"f16": 2.5e15, "bf16": 2.5e15, "fp64": 78.6e12}
PEAK_UNIT = "TFLOP/s"

# HBM bytes per launch of the distance kernels at the default workload (self-join
# n=1M d=64 eps=0.2 uniform seed 4), from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs
# of this script scaled by a streaming-copy calibration (FETCH_SIZE x 2.0, WRITE_SIZE
# x 1.0); the runs and their commands are recorded in
# profiles/selfjoin_n1M_d64_eps0.2_kernels.md.  They are not measured by this run
# (the counters need a profiler pass of their own): the line labels them as
# profile-derived, and other workloads report null.
TRAFFIC_KIB = {"f16": (395468, 4416), "fp64": (3025971, 23421)}
TRAFFIC_WORKLOAD = (1_000_000, 64, 0.2, 4)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # a join takes a few ms: 20 timed steps after 10 warm-up steps (the first
    # calls after the data generation run while the GPU clocks still ramp up)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--d", type=int, default=64)
    ap.add_argument("--eps", type=float, default=0.2)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--kernel", default="f16", choices=["f16", "bf16", "fp64"])
    ap.add_argument("--active-dims", type=int, default=-1, help="-1: chosen from the data")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-path comparison")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=25.0,
                    help="CPU baseline: the longest join that runs whole (else a sample)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-joins leg")
    ap.add_argument("--serving-only", action="store_true",
                    help="run only the serving-loop leg (prints its JSON object)")
    ap.add_argument("--serve-rates", default="200,2000",
                    help="Poisson arrival rates of the serving leg (requests/s, comma-separated)")
    ap.add_argument("--serve-seconds", type=float, default=4.0,
                    help="arrival window of the serving leg per rate and kernel")
    ap.add_argument("--serve-index", type=int, default=4_000_000,
                    help="points of the resident indexed set")
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


def roofline(stats, d, kernel, workload=None):
    """Roofline of the distance kernel of one join, from its HIP-event time.

    Executed work: 2 * d_eff flops per candidate pair, d_eff the dimensions
    (FP64 kernel, device counter) or K slots (MFMA filter: MFMAs issued x 16 x
    1024 per candidate pair) the kernel actually evaluated -- both kernels exit
    early, so most of the nominal 2 * d flops per candidate pair are never
    computed.  The nominal figure is reported beside it (frac_nominal_2d; it
    can exceed 1 for that reason)."""
    k_ms = sum(s.kernel_ms for s in stats) / len(stats)
    cand = stats[-1].candidate_pairs
    flops = 2.0 * stats[-1].pair_dims
    achieved = flops / (k_ms * 1e-3) / 1e12
    peak = PEAK[kernel] / 1e12
    traffic = None
    if workload == TRAFFIC_WORKLOAD and kernel in TRAFFIC_KIB:
        fetch, write = TRAFFIC_KIB[kernel]
        traffic = (fetch * 2.0 + write * 1.0) * 1024
    d_eff = stats[-1].pair_dims / max(cand, 1)
    nominal = 2.0 * d * cand / (k_ms * 1e-3) / 1e12
    out = {"bound": "mfma" if kernel != "fp64" else "fp64",
           "achieved": round(achieved, 3), "peak": peak,
           "unit": PEAK_UNIT, "frac": round(achieved / peak, 5), "traffic": traffic,
           "traffic_source": ("rocprofv3 --pmc FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 of this "
                              "workload (profiles/selfjoin_n1M_d64_eps0.2_kernels.md); not "
                              "measured by this run") if traffic is not None else None,
           "kernel_ms": round(k_ms, 4), "flops_per_launch": flops,
           "flops_per_unit": f"2*d_eff per candidate pair, d_eff = {d_eff:.2f}" +
           (" K slots multiplied (MFMA count x 16 x 1024 / candidate pairs)"
            if kernel != "fp64" else " dimensions evaluated"),
           "candidate_pairs_per_launch": cand,
           "nominal_2d_tflops": round(nominal, 3),
           "frac_nominal_2d": round(nominal / peak, 5)}
    return out


def sort_rate(st, n, d):
    """Achieved bytes/s of the EGO sort of one join.

    radix: the LSD passes alone (per 8-bit pass the upsweep reads the keys, 8 B
    per point, and the downsweep reads and writes keys and row indices, 2 x 12
    B per point) over the sort phase's time.
    ego_sort: the whole EGO sort as the algorithmic byte count n*d*8
    (coordinates read for the keys) + n*8 (keys written) + P*2*n*(8+4) (P
    passes over 8-byte keys with a 4-byte row) + 2*n*d*8 (the sorted
    coordinates gathered), over the grid-key, sort and index phases (the index
    phase writes the sorted operand layout and also builds the work list, so
    this rate is a lower bound)."""
    bits = st.key_dims * st.key_bits
    passes = (bits + 7) // 8
    nbytes = passes * 32 * n
    gbs = nbytes / (st.sort_ms * 1e-3) / 1e9 if st.sort_ms > 0 else None
    ego_bytes = n * d * 8 + n * 8 + passes * 2 * n * 12 + 2 * n * d * 8
    ego_ms = st.grid_ms + st.sort_ms + st.index_ms
    ego_gbs = ego_bytes / (ego_ms * 1e-3) / 1e9 if ego_ms > 0 else None
    return {"passes": passes, "key_bits": bits, "ms": round(st.sort_ms, 4),
            "algorithmic_bytes": nbytes, "GB_per_s": round(gbs, 1) if gbs else None,
            "frac_of_8TBps": round(gbs / 8000.0, 4) if gbs else None,
            "ego_sort": {"algorithmic_bytes": ego_bytes, "ms": round(ego_ms, 4),
                         "GB_per_s": round(ego_gbs, 1) if ego_gbs else None,
                         "frac_of_8TBps": round(ego_gbs / 8000.0, 4) if ego_gbs else None}}


def _near(rng, P, eps, m, lo=0.3, hi=1.0):
    """m points near m distinct rows of P: half at eps * U(lo, 0.95) (well
    within eps), half at eps * (1 +- 2^-k), k in [8, 50] (within the filter's
    rounding band, decided by the FP64 recheck)."""
    import numpy as np
    rows = rng.choice(P.shape[0], size=m, replace=False)
    w = rng.normal(size=(m, P.shape[1]))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    r = eps * rng.uniform(lo, 0.95, size=m)
    band = rng.random(m) < 0.5
    k = rng.integers(8, 51, size=m)
    r[band] = eps * (1.0 + rng.choice([-1.0, 1.0], size=int(band.sum())) * np.ldexp(1.0, -k[band]))
    return P[rows] + w * r[:, None]


def batch_leg(steps=3):
    """256 independent self-joins in one launch: sizes log-uniform in [1e3, 2e5]
    (seed 5), job k = uniform [0,1) d=64 points of seed 1000+k with 8 of its
    points replaced by near copies of other points of the job (synthetic
    neighbours, half of them within the filter's band), eps 0.2; FP16 filter
    with default and autotuned settings, and the FP64 kernel."""
    import numpy as np

    import hilbertjoin_amd as hj
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(np.log(1e3), np.log(2e5), size=256)).astype(np.int64)
    off = np.zeros(257, dtype=np.uint64)
    off[1:] = np.cumsum(sizes)
    jobs = []
    for k, n in enumerate(sizes):
        Xk = hj.gen_uniform(1000 + k, int(n), 64)
        Xk[:8] = _near(np.random.default_rng(5000 + k), Xk[8:], 0.2, 8)
        jobs.append(Xk)
    P = np.concatenate(jobs)
    del jobs
    ctx = hj.Context(0)
    dP = ctx.to_device(P)
    del P
    out = {"workload": "256 self-joins, sizes log-uniform [1e3, 2e5], d=64 uniform with 8 "
                       "planted neighbours per job, eps 0.2",
           "jobs": 256, "points": int(off[-1]),
           "pairs_total": float(sum(float(n) * (n - 1) / 2 for n in sizes))}
    rep = ctx.autotune_batch_self(dP, off, 0.2, kernel="f16")
    out["autotuned"] = rep["best"]
    out["tune_trials"] = rep["trials"]
    out["tune_base_ms"], out["tune_best_ms"] = round(rep["base_ms"], 3), round(rep["best_ms"], 3)
    counts = {}
    for name, kw in (("f16_tuned", {"kernel": "f16", "tune": True}),
                     ("f16_default", {"kernel": "f16"}), ("fp64", {"kernel": "fp64"})):
        ctx.batch_self_join_count(dP, off, 0.2, **kw)
        ts = []
        for _ in range(steps):
            t0 = time.perf_counter()
            c, st = ctx.batch_self_join_count(dP, off, 0.2, **kw)
            ts.append(time.perf_counter() - t0)
        counts[name] = c
        ms = sum(ts) / len(ts) * 1e3
        out[name] = {"ms": round(ms, 3), "kernel_ms": round(st.kernel_ms, 3),
                     "phase_ms": {"grid": round(st.grid_ms, 3), "sort": round(st.sort_ms, 3),
                                  "index": round(st.index_ms, 3), "join": round(st.join_ms, 3)},
                     "pairs_per_s": out["pairs_total"] / (ms * 1e-3),
                     "candidate_pairs": st.candidate_pairs, "band_pairs": st.band_pairs,
                     "count": int(c.sum())}
    if any(not np.array_equal(counts[k], counts["fp64"]) for k in counts):
        raise RuntimeError("batched counts differ between kernels")
    out["counts_identical_across_kernels"] = True
    dP.free()
    ctx.close()
    return out


def serving_leg(rates, seconds, n_index, kernels=("f16", "fp64")):
    """Serving loop over a resident indexed set: n_index x 32 FP64 points of a
    128-component Gaussian mixture (sigma 0.05, clipped to [0,1), seed 6),
    EGO-sorted once into HBM at eps_max = 0.10.  Requests: query sets of size
    log-uniform in [1, 65536] drawn from the same mixture (a pool of 1M further
    points of the same generator call), eps uniform in [0.02, 0.10], half
    count and half pair-list requests, Poisson arrivals at each of `rates`
    requests/s for `seconds` (seed 8).  Two mixture points lie about 0.4 apart
    at this sigma, beyond every eps, so a quarter of each request's points (up
    to 256) are replaced by synthetic near neighbours of index points (half
    well within the request's eps, half within the filter's rounding band).
    The same request stream runs through a server on the FP16 MFMA filter and
    one on the FP64 kernel; every request's count must agree between the two.
    """
    import threading

    import numpy as np

    import hilbertjoin_amd as hj
    d, pool_n, eps_max = 32, 1_000_000, 0.10
    allpts = hj.gen_mixture(6, n_index + pool_n, d, 128, 0.05)
    X, pool = allpts[:n_index], allpts[n_index:]
    ctx = hj.Context(0)
    t0 = time.perf_counter()
    index = hj.Index(ctx, X, eps_max, kernel="f16")
    build_s = time.perf_counter() - t0
    out = {"workload": f"resident {n_index} x {d} mixture (128 comps, sigma 0.05), eps_max "
                       f"{eps_max}; requests of log-uniform [1, 65536] points with a quarter "
                       f"(<= 256) planted near neighbours, eps U[0.02, 0.10], 50% pair lists, "
                       f"Poisson arrivals",
           "index_build_s": round(build_s, 4), "index": index.info(), "runs": []}
    for rate in rates:
        rng = np.random.default_rng(8)
        nreq = max(1, int(rate * seconds))
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
        sizes = np.maximum(np.minimum(np.exp(rng.uniform(0.0, np.log(65536.0), size=nreq))
                                      .astype(np.int64), 65536), 1)
        starts = rng.integers(0, pool_n - 65536, size=nreq)
        eps = rng.uniform(0.02, eps_max, size=nreq)
        want_pairs = rng.random(nreq) < 0.5
        reqs = []
        for i in range(nreq):
            Q = pool[int(starts[i]):int(starts[i]) + int(sizes[i])].copy()
            m = min(256, (int(sizes[i]) + 3) // 4)
            Q[:m] = _near(rng, X, float(eps[i]), m)
            reqs.append(Q)
        run = {"offered_rate": rate, "requests": nreq, "query_points": int(sizes.sum())}
        counts = {}
        for kernel in kernels:
            srv = hj.Server(index, workers=2, kernel=kernel, result_bytes=8 << 30)
            for p in (False, True):  # warm the workers' buffers
                srv.wait(srv.submit(reqs[0], eps_max, pairs=p))
            srv.reset_stats()
            tickets = [None] * nreq
            ev = [threading.Event() for _ in range(nreq)]

            def submitter():
                base = time.perf_counter()
                for i in range(nreq):
                    delay = arrivals[i] - (time.perf_counter() - base)
                    if delay > 0:
                        time.sleep(delay)
                    tickets[i] = srv.submit(reqs[i], float(eps[i]), pairs=bool(want_pairs[i]))
                    ev[i].set()

            th = threading.Thread(target=submitter)
            t_start = time.perf_counter()
            th.start()
            got, npairs = [], 0
            for i in range(nreq):
                ev[i].wait()
                r = srv.wait(tickets[i])
                if r.status != 0:
                    raise RuntimeError(f"serving request {i} failed with status {r.status}")
                got.append(r.count)
                if r.colidx is not None:
                    npairs += r.colidx.size
            th.join()
            wall = time.perf_counter() - t_start
            st = srv.stats()
            srv.close()
            counts[kernel] = got
            run[kernel] = {"p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                           "mean_ms": round(st["mean_ms"], 3), "max_ms": round(st["max_ms"], 3),
                           "requests_per_s": round(st["requests_per_s"], 2),
                           "query_points_per_s": round(float(sizes.sum()) / wall, 1),
                           "batches": st["batches"],
                           "mean_batch_requests": round(st["mean_batch_requests"], 2),
                           "pairs_returned": int(npairs), "pairs_counted": int(sum(got))}
        if len(kernels) > 1 and any(counts[k] != counts[kernels[0]] for k in kernels[1:]):
            raise RuntimeError("serving counts differ between kernels")
        run["counts_identical_across_kernels"] = len(kernels) > 1
        out["runs"].append(run)
    index.close()
    ctx.close()
    return out


def cpu_baseline(X, eps, target_s):
    """The oracle's CPU EGO join (restatement of the reference algorithm: EGO
    sort, per-point stripe ranges, 8x8 blocks with an early exit every 4
    dimensions; C, OpenMP, -O3 -march=x86-64-v3) on the host cores.  A short
    sampled run (the join loop on every 4096-th 8-point block) estimates the
    full join; if that fits in target_s the whole join runs and nothing is
    extrapolated, otherwise the join loop runs on every stride-th block and
    its time is scaled by the stride."""
    import oracle
    n = X.shape[0]
    threads = oracle.max_threads()
    stride = 4096
    c, t, ev = oracle.ego_self_count(X, eps, active_dims=3, stride=stride)
    est = t[0] + t[1] + t[2] * stride
    if est > target_s and t[2] > 0:
        stride = max(1, int(stride * t[2] / max(target_s - t[0] - t[1], 1.0)))
    else:
        stride = 1
    c, t, ev = oracle.ego_self_count(X, eps, active_dims=3, stride=stride)
    join_full = t[2] * stride
    total = t[0] + t[1] + join_full
    pairs = n * (n - 1) / 2
    sample = (f"the whole CPU EGO join of the same {n} points (sort {t[0]:.2f} s, index "
              f"{t[1]:.2f} s, join loop {t[2]:.2f} s, {ev} pair evaluations, count {c})"
              if stride == 1 else
              f"CPU EGO join of the same {n} points (sort {t[0]:.2f} s + index {t[1]:.2f} s in "
              f"full; join loop on 1/{stride} of the 8-point blocks, {t[2]:.2f} s, {ev} pair "
              f"evaluations, extrapolated x{stride})")
    eps_rate = ev / t[2] if t[2] > 0 else None
    return {"value": pairs / total, "unit": UNIT, "cores": threads, "kind": "port",
            "sample": sample, "join_s": round(total, 3), "extrapolated": stride > 1,
            "evaluations": ev, "evaluations_per_s": eps_rate,
            "evaluations_per_s_per_core": eps_rate / threads if eps_rate else None,
            "build": "gcc -O3 -march=x86-64-v3 -ffp-contract=off -fopenmp (oracle/Makefile)"}


def main():
    a = parse()
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "the join runs on exactly one GPU; --gpus must be 1",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.serving_only:
        print(json.dumps(serving_leg([float(r) for r in a.serve_rates.split(",")],
                                     a.serve_seconds, a.serve_index)))
        return
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
        "step_ms": {"min": round(1e3 * min(walls), 4),
                    "median": round(1e3 * sorted(walls)[len(walls) // 2], 4),
                    "max": round(1e3 * max(walls), 4)},
        "executed_pairs_per_s": st.candidate_pairs / (ms * 1e-3),
        "candidate_pairs": st.candidate_pairs, "load_fraction": round(st.load_fraction, 6),
        "tile_pairs": st.tile_pairs, "band_pairs": st.band_pairs,
        "phase_ms": {"grid": round(st.grid_ms, 4), "sort": round(st.sort_ms, 4),
                     "index": round(st.index_ms, 4), "join": round(st.join_ms, 4),
                     "kernel": round(st.kernel_ms, 4)},
        "roofline": roofline(stats, a.d, a.kernel, (a.n, a.d, a.eps, a.seed)),
        "sort": sort_rate(st, a.n, a.d),
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
                            "roofline": roofline(s64, a.d, "fp64", (a.n, a.d, a.eps, a.seed))}
        out["speedup_vs_fp64"] = round(ms64 / ms, 3)
    dev.free()
    ctx.close()
    if not a.no_batch:
        out["batch"] = batch_leg()
    out["serving_p99_ms"] = None
    if not a.no_serving:
        sv = serving_leg([float(r) for r in a.serve_rates.split(",")], a.serve_seconds,
                         a.serve_index)
        out["serving"] = sv
        first = sv["runs"][0]
        out["serving_p99_ms"] = first[a.kernel if a.kernel in first else "fp64"]["p99_ms"]
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(X, a.eps, a.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
