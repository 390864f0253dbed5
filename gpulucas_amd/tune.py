"""Measure the fastest tiling per signal length and write the shape table.

For each length N the engine's admissible tilings (A x B radix split, tile
width C, workgroup size) are timed on the GPU with `lucas_autotune_shape`;
the winners go to gpulucas_amd/csrc/tuned_shapes.inc, which the planner reads
before its cost model (rebuild the library afterwards).

    python -m gpulucas_amd.tune --out tune.json            # measure the standard lengths
    python -m gpulucas_amd.tune --range --out more.json    # every other length of p in [1e5, 4e7]
    python -m gpulucas_amd.tune --emit tune.json more.json # write the .inc
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
INC = os.path.join(HERE, "csrc", "tuned_shapes.inc")


def standard_lengths(serving: bool = True):
    """(p, N) pairs of the standard workloads: config shapes, batch lengths and
    (optionally) the serving workload's lengths."""
    from gpulucas_amd import lucas, workloads as wl

    pairs = {4608: 86243, 65536: 1257787, 163840: 2976221, 49152: 859433}
    for p in wl.batch_primes() + wl.near3m_primes():
        pairs.setdefault(lucas.default_N(p), p)
    if serving:
        for _, p, _ in wl.serving_jobs(count=64):
            pairs.setdefault(lucas.default_N(p), p)
    return sorted((p, N) for N, p in pairs.items())


def tuned_entries():
    """Entries of the shape table tuned_shapes.inc as
    (N, A, C, threads, pref, wave_A, wave_B, p), p the exponent measured."""
    import re

    rows = []
    with open(INC) as f:
        for line in f:
            m = re.match(r"\{(-?\d+), (-?\d+), (-?\d+), (-?\d+), (-?\d+), (-?\d+), (-?\d+)\},.* at p=(\d+)", line)
            if m:
                rows.append(tuple(int(x) for x in m.groups()))
    return rows


def range_lengths(lo: int = 100_000, hi: int = 40_000_000, step: float = 0.002):
    """(p, N) for every signal length the default rule picks for primes in
    [lo, hi] (exponents sampled at relative steps of `step`, each rounded up
    to the next prime): the lengths a served job can run at."""
    from gpulucas_amd import lucas, workloads as wl

    pairs = {}
    p = lo
    while p < hi:
        q = p
        while not wl.is_prime(q):
            q += 1
        pairs.setdefault(lucas.default_N(q), q)
        p = int(p * (1 + step)) + 1
    return sorted((p, N) for N, p in pairs.items())


def measure(pairs, iters, max_candidates, out):
    from gpulucas_amd import lucas

    results = {}
    for p, N in pairs:
        t0 = time.time()
        info, ms = lucas.autotune_shape(p, N, iters=iters, max_candidates=max_candidates)
        base = lucas.plan(p, N)
        results[str(N)] = {"p": p, "N": N, "A": info["A"], "C": info["C"], "threads": info["threads"],
                           "radix_pref": info["radix_pref"], "wave_A": info["wave_A"], "wave_B": info["wave_B"],
                           "ms_per_iter": ms,
                           "model_A": base["A"], "model_C": base["C"]}
        print(f"N={N} p={p}: A={info['A']} C={info['C']} T={info['threads']} R{info['radix_pref']} "
              f"W{info['wave_A']}/{info['wave_B']} "
              f"{ms * 1e3:.2f} us/iter "
              f"({time.time() - t0:.1f}s)", flush=True)
        with open(out, "w") as f:
            json.dump(results, f, indent=1)
    return results


def emit(paths):
    table = {}
    for path in paths:
        with open(path) as f:
            table.update(json.load(f))
    rows = sorted((int(v["N"]), int(v["A"]), int(v["C"]), int(v["threads"]), int(v.get("radix_pref", 8)),
                   int(v.get("wave_A", -1)), int(v.get("wave_B", -1)), float(v["ms_per_iter"]), int(v["p"]))
                  for v in table.values())
    with open(INC, "w") as f:
        f.write("// Generated by gpulucas_amd/tune.py from GPU measurements (MI355X):\n")
        f.write("// {N, A, C, threads, radix preference, column / row wave factor}\n")
        for N, A, C, T, pref, wa, wb, ms, p in rows:
            f.write(f"{{{N}, {A}, {C}, {T}, {pref}, {wa}, {wb}}},  // {ms * 1e3:.2f} us/iteration at p={p}\n")
    print(f"wrote {len(rows)} entries to {INC}")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--emit", nargs="*", default=None)
    ap.add_argument("--iters", type=int, default=300)
    ap.add_argument("--max-candidates", type=int, default=250)
    ap.add_argument("--lengths", default=None, help="comma list of p:N pairs")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--range", action="store_true",
                    help="every length the default rule picks for p in [1e5, 4e7], not in the table yet")
    a = ap.parse_args(argv)
    if a.emit is not None:
        emit(a.emit)
        return
    if a.lengths:
        pairs = [tuple(int(x) for x in s.split(":")) for s in a.lengths.split(",")]
    elif a.range:
        known = {N for _, N in standard_lengths()}
        pairs = [(p, N) for p, N in range_lengths() if N not in known]
    else:
        pairs = standard_lengths(serving=not a.no_serving)
    measure(pairs, a.iters, a.max_candidates, a.out or "tune.json")


if __name__ == "__main__":
    sys.exit(main())
