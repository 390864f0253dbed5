#!/usr/bin/env python3
"""Generate the GMP Lucas-Lehmer Res64 fixtures in tests/golden/gmp_res64.json.

Each entry is the low 64 bits of the canonical s_k mod (2^p - 1) after k
squarings from s_0 = 4, computed by the GMP oracle (oracle/gmp_ll, built by
``make -C oracle``).  Runs are independent and spread over a process pool;
each run's lines are kept in a scratch folder (a fresh temporary folder
unless --scratch names one), so a generation interrupted with --scratch
resumes where it stopped.

Usage:
    python tests/golden/make_gmp_fixtures.py [--jobs 6] [--scratch DIR] [--only NAME]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from gpulucas_amd import workloads as wl  # noqa: E402

GMP_LL = os.path.join(ROOT, "oracle", "gmp_ll")
OUT = os.path.join(HERE, "gmp_res64.json")

# Small composite-exponent primes (full-length runs, non-zero residues).
SMALL_COMPOSITE_MP = (1009, 2003, 4001, 9973, 10007, 20011, 30011, 44501, 50021)


def plan():
    """name -> list of (p, [iterations])."""
    runs = {}
    # Reference configuration shapes.
    runs["c1_p86243"] = [(86243, [1000, 10000, 86241])]
    runs["c2_p1257787"] = [(1257787, [1000, 10000] + [100000 * m for m in range(1, 13)] + [1257785])]
    runs["c3_p2976221"] = [(2976221, [1000, 5000, 20000, 100000])]
    runs["c3_near3m"] = [(p, [1000, 5000, 20000]) for p in wl.near3m_primes()]
    runs["c4_batch64"] = [(p, [1000, 10000]) for p in wl.batch_primes()]
    # Small known answers: every Mersenne-prime exponent in [127, 216091] to the end.
    small = [p for p in wl.MERSENNE_PRIME_EXPONENTS if 127 <= p <= 216091]
    runs["small_primes_full"] = [(p, sorted({1, 2, 5, 100, p - 2})) for p in small]
    runs["small_composite_full"] = [(p, [1, 5, 100, 1000, p - 2]) for p in SMALL_COMPOSITE_MP]
    runs["mid_p859433"] = [(859433, [1000, 50000, 859431])]
    # Long transforms (N = 2^19 .. 2.3M): the top of the serving range.
    runs["large"] = [(9600037, [1000, 5000]), (20996011, [100, 2000]), (37156667, [100, 2000]),
                     (37732501, [2000])]
    # The longest lengths: M48 (N = 3670016) and, past the ordinary tilings
    # (wide tiles, N up to 8064000), a depth-4 exponent and M52.
    runs["longest"] = [(57885161, [100]), (100000007, [100]), (136279841, [100])]
    # The serving workload at its stated size: the first 48 Poisson jobs of
    # seed 5 (p log-uniform in [1e5, 4e7], K uniform in [5000, 50000]), each
    # at its own full K, plus an early checkpoint.
    runs["c5_serving"] = [(p, sorted({1000, k})) for _, p, k in wl.serving_jobs(seed=5, count=48)]
    return runs


def run_one(p, iters, scratch):
    path = os.path.join(scratch, f"p{p}.txt")
    done = {}
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                k, h, z = line.split()
                done[int(k)] = (h, int(z))
    todo = [k for k in iters if k not in done]
    if todo:
        cmd = [GMP_LL, str(p)] + [str(k) for k in sorted(set(iters))]
        with open(path + ".tmp", "w") as f:
            subprocess.run(cmd, stdout=f, check=True)
        os.replace(path + ".tmp", path)
        with open(path) as f:
            for line in f:
                k, h, z = line.split()
                done[int(k)] = (h, int(z))
    return p, {str(k): done[k][0] for k in iters}, {str(k): done[k][1] for k in iters}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=6)
    ap.add_argument("--scratch", default=None, help="resumable scratch folder (default: a temporary one)")
    ap.add_argument("--only", default=None)
    ap.add_argument("--out", default=OUT)
    args = ap.parse_args()
    out_path = args.out
    if not os.path.exists(GMP_LL):
        sys.exit("build the oracle first: make -C oracle")
    if args.scratch is None:
        args.scratch = tempfile.mkdtemp(prefix="gmp_fixtures_")
    os.makedirs(args.scratch, exist_ok=True)
    runs = plan()
    if args.only:
        runs = {k: v for k, v in runs.items() if k in args.only.split(",")}
    existing = {}
    if os.path.exists(out_path):
        with open(out_path) as f:
            existing = json.load(f)
    result = dict(existing)
    # Longest runs first so the pool drains evenly.
    tasks = [(name, p, its) for name, lst in runs.items() for (p, its) in lst]
    tasks.sort(key=lambda t: -(t[1] ** 1.6) * max(t[2]))
    with cf.ProcessPoolExecutor(max_workers=args.jobs) as ex:
        futs = {ex.submit(run_one, p, its, args.scratch): (name, p) for name, p, its in tasks}
        for fut in cf.as_completed(futs):
            name, p = futs[fut]
            _, res, zero = fut.result()
            group = result.setdefault(name, {})
            group[str(p)] = {"res64": res, "zero": zero}
            print(f"done {name} p={p}", flush=True)
            with open(out_path + ".tmp", "w") as f:
                json.dump(result, f, indent=1, sort_keys=True)
            os.replace(out_path + ".tmp", out_path)
    result["_meta"] = {
        "generator": "tests/golden/make_gmp_fixtures.py",
        "oracle": "oracle/gmp_ll.c (GMP mpz squaring, Mersenne fold)",
        "convention": "res64 of s_k after k squarings from s_0=4; zero=1 when s_k == 0 mod 2^p-1",
    }
    with open(out_path + ".tmp", "w") as f:
        json.dump(result, f, indent=1, sort_keys=True)
    os.replace(out_path + ".tmp", out_path)


if __name__ == "__main__":
    main()
