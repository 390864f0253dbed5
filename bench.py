#!/usr/bin/env python3
"""Lucas-Lehmer iteration benchmark on one MI355X.

One step = one LL iteration s <- s^2 - 2 mod 2^p - 1 of the headline
workload: p = 2976221, N = 163840 = 2^15*5 (FP64 IBDWT, state resident in
HBM).  Prints one JSON line with ms per iteration, the roofline of the
dominant kernel, the batched-vs-plain throughput ratio and the GMP CPU
baseline timed on this host.

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
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def headline(lucas, steps, warmup):
    ctx = lucas.Context(P_HEAD, N_HEAD)
    info = ctx.info
    log(f"plan {info}")
    ctx.step(max(warmup, 1))  # untimed: fills s to p bits, builds the graph
    ctx.max_roundoff(reset=True)
    t0 = time.perf_counter()
    ctx.step(steps)  # ends with a device synchronisation
    t1 = time.perf_counter()
    ms = 1e3 * (t1 - t0) / steps
    err = ctx.max_roundoff()
    col_us, row_us = ctx.time_kernels(200)
    ctx.close()
    return ms, err, col_us, row_us, info


def batched_ratio(lucas, wl, iters, count):
    """Aggregate iterations/s: all exponents in one batch vs one at a time."""
    ps = wl.batch_primes(count=count)
    b = lucas.Batch(ps)
    b.step(2)
    ms_batch = b.time_iterations(iters)
    b.close()
    batched_its = count * 1e3 / ms_batch
    total_ms = 0.0
    for p in ps:
        c = lucas.Context(p, 0)
        c.step(2)
        total_ms += c.time_iterations(iters) * iters
        c.close()
    plain_its = count * iters * 1e3 / total_ms
    return batched_its, plain_its


def cpu_baseline(seconds_target):
    import oracle

    # calibrate on a short run, then time a bounded sample (~seconds_target s)
    ms = oracle.gmp_time_ms(P_HEAD, 64, 20)
    iters = max(50, int(seconds_target * 1e3 / ms))
    ms = oracle.gmp_time_ms(P_HEAD, 0, iters)
    return {"value": round(ms, 4), "unit": "ms/LL-iter", "cores": 1, "kind": "port",
            "sample": f"GMP {oracle.gmp_version()} LL squaring+fold at p={P_HEAD}, {iters} iterations "
                      f"after 64 warm-up, 1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--batch-iters", type=int, default=300)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); the engine runs one "
              "process per GPU and exponents shard across GPUs without communication.", file=sys.stderr)
        sys.exit(2)

    from gpulucas_amd import lucas, workloads as wl

    ms, err, col_us, row_us, info = headline(lucas, a.steps, a.warmup)
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
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None,
                     "kernel_us": {"lucas_col_kernel": round(col_us, 3), "lucas_row_kernel": round(row_us, 3)},
                     "algorithmic_bytes_per_launch": kern_bytes},
        "fp64_tflops": round(5 * N * __import__("math").log2(N) / (ms * 1e-3) / 1e12, 4),
    }
    if not a.no_batch:
        bi, pi = batched_ratio(lucas, wl, a.batch_iters, 64)
        out["batched"] = {"workload": "64 primes in [1e6,5e6] (seed 4), default N each",
                          "iters_per_s": round(bi, 1), "plain_iters_per_s": round(pi, 1),
                          "ratio": round(bi / pi, 3)}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
