#!/usr/bin/env python3
"""Lucas-Lehmer iteration benchmark on one MI355X.

One step = one LL iteration s <- s^2 - 2 mod 2^p - 1 of the headline
workload: p = 2976221, N = 163840 = 2^15*5 (FP64 IBDWT, state resident in
HBM).  Prints one JSON line with ms per iteration, the roofline of the
dominant kernel (with its HBM traffic from the committed rocprofv3 counter
summary), the per-iteration roofline of a 2^21-point transform
(p = 37156667), the batched-vs-plain throughput ratios, the serving
p50/p99 latency and capacity, and the GMP CPU baseline timed on this host.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/LL-iter at p=2976221 N=160K; batched iters/s vs plain; p50/p99 latency"
P_HEAD, N_HEAD = 2976221, 163840
P_LONG, N_LONG = 37156667, 2097152  # Mersenne prime, N = 2^21
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_times(ctx, iters):
    """Mean duration of each kernel in the steady-state iteration.

    The iteration time t_iter comes from `iters` graph-launched iterations
    (graphs built beforehand), the same launches a step makes.  HIP events
    recorded after every launch of the row/col sequence on the engine's
    stream give each kernel's interval; every interval also holds one
    marker's own cost m, the same for both, so m = (col_i + row_i - t_iter)/2
    and each kernel is its interval less m.  Neither figure depends on
    --steps.  Returns (col_us, row_us, t_iter_us, raw)."""
    t_iter = 1e3 * ctx.time_iterations(iters)
    col_i, row_i = ctx.time_kernels(200)
    m = max(0.0, (col_i + row_i - t_iter) / 2)
    col, row = col_i - m, row_i - m
    raw = {"iteration_us": round(t_iter, 3), "col_interval_us": round(col_i, 3),
           "row_interval_us": round(row_i, 3), "marker_us": round(m, 3), "graph_iterations": iters,
           "method": "graph-launched iteration time; event-timed launch intervals less the per-marker cost "
                     "(col_i + row_i - t_iter) / 2"}
    return col, row, t_iter, raw


def headline(lucas, steps, warmup):
    ctx = lucas.Context(P_HEAD, N_HEAD)
    info = ctx.info
    log(f"plan {info}")
    ctx.step(max(warmup, 1))  # untimed warm-up iterations: s fills all p bits
    # set-up: build the replay graph of a `steps`-iteration step and launch it
    # once, restoring the state (a graph's first launch costs ~30 us extra)
    ctx.prepare(steps)
    ctx.max_roundoff(reset=True)
    t0 = time.perf_counter()
    ctx.step(steps)  # ends with a device synchronisation
    t1 = time.perf_counter()
    ms = 1e3 * (t1 - t0) / steps
    err = ctx.max_roundoff()
    bits = ctx.product_bits()
    col_us, row_us, t_iter, raw = kernel_times(ctx, 2000)
    ctx.close()
    return ms, err, bits, col_us, row_us, info, raw


def long_transform(lucas, p, N, steps, warmup):
    """ms/iter at a long transform (N >= 2^21), where the iteration should be
    HBM-bound: achieved = 32N algorithmic bytes per iteration / iteration time."""
    with lucas.Context(p, N) as ctx:
        info = ctx.info
        ctx.step(max(warmup, 1))
        ctx.prepare(steps)
        ctx.max_roundoff(reset=True)
        t0 = time.perf_counter()
        ctx.step(steps)
        t1 = time.perf_counter()
        err = ctx.max_roundoff()
        us = 1e6 * (t1 - t0) / steps
        col_us, row_us, _, _ = kernel_times(ctx, 1000)
    gbs = 32 * N / (us * 1e-6) / 1e9
    return {"workload": f"LL iteration p={p} N={N}, FP64 IBDWT", "p": p, "N": N,
            "A": info["A"], "B": info["B"], "C": info["C"], "steps": steps,
            "us_per_iter": round(us, 3), "iters_per_s": round(1e6 / us, 1), "max_roundoff": round(err, 5),
            "roofline_iteration": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                                   "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5),
                                   "algorithmic_bytes_per_iter": 32 * N},
            "kernel_us": {"lucas_col_kernel": round(col_us, 3), "lucas_row_kernel": round(row_us, 3)}}


def batched_ratio(lucas, wl, iters, count, seed=4, lo=1_000_000, hi=5_000_000):
    """Aggregate iterations/s: all exponents in one batch vs one at a time.
    Each side first runs `iters` untimed iterations, which builds its replay
    graphs, so neither timed region includes graph instantiation."""
    ps = wl.batch_primes(seed=seed, count=count, lo=lo, hi=hi)
    b = lucas.Batch(ps)
    b.step(iters)
    ms_batch = b.time_iterations(iters)
    b.close()
    batched_its = count * 1e3 / ms_batch
    total_ms = 0.0
    for p in ps:
        c = lucas.Context(p, 0)
        c.step(iters)
        total_ms += c.time_iterations(iters) * iters
        c.close()
    plain_its = count * iters * 1e3 / total_ms
    return batched_its, plain_its


def measured_traffic(info):
    """HBM bytes per launch of each kernel from the committed counter summary
    (profiles/traffic_p2976221_N163840.md), if it was taken with this plan."""
    path = os.path.join(ROOT, "profiles", "traffic_p2976221_N163840.md")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if line.startswith("traffic:"):
            k, v = line[len("traffic:"):].split("=", 1)
            vals[k.strip()] = v.strip()
    shape = (f"A={info['A']} B={info['B']} C={info['C']} T={info['threads']} R={info['radix_pref']} "
             f"W={info['wave_A']}/{info['wave_B']} L={info['line']}")
    if vals.get("plan") != shape:
        return None
    return {k: int(v) for k, v in vals.items() if k != "plan"}


def run_jobs(lucas, jobs, scale):
    """Submit (t, p, K) jobs to a continuous-batching server at times
    t * scale (all at once for scale 0); returns (results, stats)."""
    ids = {}
    with lucas.Server(chunk_iters=256) as srv:
        t0 = time.perf_counter()
        for t, p, k in jobs:
            dt = t * scale - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids[srv.submit(p, k)] = (p, k)
        srv.drain(timeout_s=900)
        res = [srv.poll(j) for j in ids]
        st = srv.stats()
    return res, st


def serving(lucas, wl, count, load):
    """The serving workload at its stated size: the first `count` Poisson
    jobs of seed 5 (p log-uniform in [1e5, 4e7], K uniform in [5000, 50000],
    each at its full K).  The server's capacity is measured first, as the
    algorithmic HBM bytes (32 N per iteration) per second it sustains with
    every job submitted at once; the Poisson arrivals are then spaced so
    that the offered load is `load` x that capacity."""
    jobs = wl.serving_jobs(seed=5, count=count)
    ns = [lucas.default_N(p) for _, p, _ in jobs]
    work = sum(32.0 * n * k for n, (_, _, k) in zip(ns, jobs))
    log(f"serving: {count} jobs, {sum(k for _, _, k in jobs)} iterations, {work / 1e12:.2f} TB algorithmic")
    res, st = run_jobs(lucas, jobs, 0.0)
    span_cap = max(r["finish_s"] for r in res) - min(r["submit_s"] for r in res)
    capacity = work / span_cap  # bytes/s
    log(f"serving capacity: {span_cap:.2f} s for all jobs at once = {capacity / 1e12:.3f} TB/s")
    span_arr = work / (load * capacity)
    res, st = run_jobs(lucas, jobs, span_arr / jobs[-1][0])
    lat = sorted(r["finish_s"] - r["submit_s"] for r in res)
    span = max(r["finish_s"] for r in res) - min(r["submit_s"] for r in res)

    def pct(q):
        return lat[min(len(lat) - 1, int(round(q * (len(lat) - 1))))]

    offered = load * capacity
    return {"workload": f"{count} Poisson jobs (seed 5), p log-uniform [1e5,4e7], N {min(ns)}..{max(ns)}, "
                        f"K in [5000,50000] at full size",
            "jobs": len(res), "iterations": sum(k for _, _, k in jobs),
            "p50_s": round(pct(0.5), 4), "p99_s": round(pct(0.99), 4), "max_s": round(lat[-1], 4),
            "iters_per_s": round(sum(k for _, _, k in jobs) / span, 1),
            "capacity": {"all_at_once_s": round(span_cap, 3), "algorithmic_TBps": round(capacity / 1e12, 4)},
            "offered_load": {"of_capacity": load, "algorithmic_TBps": round(offered / 1e12, 4),
                             "of_hbm_peak": round(offered / (HBM_PEAK_GBS * 1e9), 4)},
            "arrival_span_s": round(span_arr, 3), "peak_jobs": st["peak_jobs"],
            "busy_s": round(st["busy_s"], 3), "span_s": round(span, 3),
            "max_roundoff": round(max(r["max_roundoff"] for r in res), 4),
            "all_ok": all(r["status"] == 0 for r in res)}


def serving_capacity(lucas, wl, plain_its, iters=2000, count=64, seed=6, lo=100_000, hi=1_200_000):
    """Aggregate iterations/s of the continuous-batching server when `count`
    jobs of N <= 64K arrive at once (the same exponents as `batched_small`),
    against `plain_its`, the rate of running them one after another."""
    ps = wl.batch_primes(seed=seed, count=count, lo=lo, hi=hi)
    with lucas.Server(chunk_iters=256) as srv:
        ids = [srv.submit(p, iters) for p in ps]
        srv.drain(timeout_s=600)
        res = [srv.poll(j) for j in ids]
    span = max(r["finish_s"] for r in res) - min(r["submit_s"] for r in res)
    its = count * iters / span
    return {"workload": f"{count} jobs of {iters} iterations submitted at once, primes in [1e5,1.2e6] "
                        f"(seed {seed}), N <= 64K",
            "iters_per_s": round(its, 1), "plain_iters_per_s": round(plain_its, 1),
            "ratio": round(its / plain_its, 3), "all_ok": all(r["status"] == 0 for r in res)}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(seconds_target):
    """GMP Lucas-Lehmer (the oracle's squaring + Mersenne fold) at the
    headline exponent on this host's cores, as child processes of the oracle
    binary, each pinned to one core: one process alone (ms/iteration), then
    one per core at once (aggregate iterations/s).  Runs before any GPU
    call; bounded to about `seconds_target` seconds per phase."""
    import subprocess

    import oracle

    gmp = os.path.join(ROOT, "oracle", "gmp_ll")
    if not os.path.exists(gmp):
        oracle.build()
    cores = sorted(os.sched_getaffinity(0))[:16]  # the box's CPU share for one GPU is 16

    def run(core, warm, iters):
        return subprocess.Popen([gmp, "--time", str(P_HEAD), str(warm), str(iters)], stdout=subprocess.PIPE,
                                text=True, preexec_fn=lambda: os.sched_setaffinity(0, {core}))

    # calibrate on a short run, then a bounded sample on one core
    ms = float(run(cores[0], 64, 20).communicate()[0])
    iters = max(50, int(seconds_target * 1e3 / ms))
    ms1 = float(run(cores[0], 64, iters).communicate()[0])
    # every core at once, same sample each
    procs = [run(c, 64, iters) for c in cores]
    per = [float(p.communicate()[0]) for p in procs]
    agg = sum(1e3 / m for m in per)
    return {"value": round(ms1, 4), "unit": "ms/LL-iter", "cores": 1, "kind": "port",
            "sample": f"GMP {oracle.gmp_version()} LL squaring+fold at p={P_HEAD}, {iters} iterations after "
                      f"64 warm-up, 1 pinned process",
            "cpu_model": cpu_model(),
            "aggregate": {"cores": len(cores), "iters_per_s": round(agg, 2),
                          "ms_per_iter_per_core": round(sum(per) / len(per), 4),
                          "sample": f"{len(cores)} pinned processes, one per core, {iters} iterations each"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--batch-iters", type=int, default=300)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-long", action="store_true")
    ap.add_argument("--long-steps", type=int, default=3000)
    ap.add_argument("--serve-jobs", type=int, default=48)
    ap.add_argument("--profile-run", action="store_true",
                    help="only the headline iterations (for rocprofv3 kernel-trace / counter passes)")
    ap.add_argument("--p", type=int, default=P_HEAD, help="exponent for --profile-run")
    ap.add_argument("--N", type=int, default=N_HEAD, help="signal length for --profile-run (0: default)")
    ap.add_argument("--line", default="", metavar="A:K[:KIND]",
                    help="--profile-run on the line kernels with A rows and K own columns per column workgroup "
                         "(KIND: lucas_plan_info.line, default 1)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only a device-to-device copy of BYTES (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); the engine runs one "
              "process per GPU and exponents shard across GPUs without communication.", file=sys.stderr)
        sys.exit(2)

    # the CPU baseline runs first, before the engine's library is loaded or
    # any GPU call is made (its workers are child processes pinned to cores)
    cpu = None
    if not (a.no_cpu or a.copy_calibration or a.profile_run):
        cpu = cpu_baseline(a.cpu_seconds)
        log(f"cpu baseline {cpu}")

    from gpulucas_amd import lucas, workloads as wl

    if a.copy_calibration:
        us = lucas.device_copy_bench(a.copy_calibration, reps=20)
        print(json.dumps({"copy_bytes": a.copy_calibration, "us_per_copy": us,
                          "GBps": 2 * a.copy_calibration / us / 1e3}), flush=True)
        return
    if a.profile_run:
        plan = None
        if a.line:
            f = [int(v) for v in a.line.split(":")]
            plan = {"p": a.p, "N": a.N, "A": f[0], "C": f[1], "line": f[2] if len(f) > 2 else 1}
        with lucas.Context(a.p, a.N, plan=plan) as ctx:
            ctx.step(a.warmup + a.steps)
            print(json.dumps({"profile_run": ctx.info, "max_roundoff": ctx.max_roundoff()}), flush=True)
        return

    ms, err, bits, col_us, row_us, info, raw = headline(lucas, a.steps, a.warmup)
    N = N_HEAD
    kern_bytes = 16 * N  # one read + one write of the 8N-byte state per kernel
    dom_name, dom_us = ("lucas_col_kernel", col_us) if col_us >= row_us else ("lucas_row_kernel", row_us)
    achieved = kern_bytes / (dom_us * 1e-6) / 1e9
    out = {
        "metric": METRIC,
        "value": round(ms, 6),
        "unit": "ms/iter",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 6),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (LL sequence from s0=4)",
        "config": {"workload": f"LL iteration p={P_HEAD} N={N} (2^15*5), FP64 IBDWT",
                   "p": P_HEAD, "N": N, "A": info["A"], "B": info["B"], "C": info["C"],
                   "parallelism": "single exponent, 1 GPU"},
        "iters_per_s": round(1e3 / ms, 1),
        "max_roundoff": round(err, 5),
        "product_bits": {"max_bit_length": bits[0], "max_leading_zeros": bits[1]},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None,
                     "kernel_us": {"lucas_col_kernel": round(col_us, 3), "lucas_row_kernel": round(row_us, 3)},
                     "kernel_timing": raw,
                     "algorithmic_bytes_per_launch": kern_bytes},
        "fp64_tflops": round(5 * N * __import__("math").log2(N) / (ms * 1e-3) / 1e12, 4),
    }
    tr = measured_traffic(info)
    if tr:
        key = "lucas_col_kernel" if dom_name == "lucas_col_kernel" else "lucas_row_kernel"
        out["roofline"]["traffic"] = tr.get(key)
        out["roofline"]["traffic_by_kernel"] = tr
    if not a.no_long:
        out["long_transform"] = long_transform(lucas, P_LONG, N_LONG, a.long_steps, 100)
    if not a.no_batch:
        bi, pi = batched_ratio(lucas, wl, a.batch_iters, 64)
        out["batched"] = {"workload": "64 primes in [1e6,5e6] (seed 4), default N each",
                          "iters_per_s": round(bi, 1), "plain_iters_per_s": round(pi, 1),
                          "ratio": round(bi / pi, 3)}
        bi, pi = batched_ratio(lucas, wl, a.batch_iters, 64, seed=6, lo=100_000, hi=1_200_000)
        out["batched_small"] = {"workload": "64 primes in [1e5,1.2e6] (seed 6), default N each (N <= 64K)",
                                "iters_per_s": round(bi, 1), "plain_iters_per_s": round(pi, 1),
                                "ratio": round(bi / pi, 3)}
        if not a.no_serve:
            out["serving_capacity"] = serving_capacity(lucas, wl, pi)
    if not a.no_serve:
        out["serving"] = serving(lucas, wl, a.serve_jobs, 0.8)
    if cpu:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
