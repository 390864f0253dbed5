"""Measure the fastest tiling per signal length and write the shape table.

For each length N the engine's admissible tilings (A x B radix split, tile
width C, workgroup size) are timed on the GPU with `lucas_autotune_shape`;
the winners go to gpulucas_amd/csrc/tuned_shapes.inc, which the planner reads
before its cost model (rebuild the library afterwards).

    python -m gpulucas_amd.tune --out tune.json            # measure the standard lengths
    python -m gpulucas_amd.tune --range --out more.json    # every other length of p in [1e5, 4e7]
    python -m gpulucas_amd.tune --emit tune.json more.json # write the .inc
    python -m gpulucas_amd.tune --line --out line.json     # line kernels vs the table's tilings
    python -m gpulucas_amd.tune --emit-line line.json      # add the line shapes that win to the .inc

A length may hold two entries: the tiling of the general kernels (which
batches always use) and, where the line kernels measured faster for a single
exponent, a line entry (last field 1, or 2 with non-temporal spectrum loads
and stores; 3 and 4 the same with every row split over two waves), which
single exponents then use.
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
    (N, A, C, threads, pref, wave_A, wave_B, line, p), p the exponent measured."""
    import re

    rows = []
    with open(INC) as f:
        for line in f:
            m = re.match(r"\{(-?\d+), (-?\d+), (-?\d+), (-?\d+), (-?\d+), (-?\d+), (-?\d+)(?:, (\d+))?\},.* at p=(\d+)",
                         line)
            if m:
                g = m.groups()
                rows.append(tuple(int(x) for x in g[:7]) + (int(g[7] or 0), int(g[8])))
    return rows


def _comment_us(line: str) -> float | None:
    import re

    m = re.search(r"// ([0-9.]+) us/iteration", line)
    return float(m.group(1)) if m else None


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


def _write_inc(rows):
    """rows: (N, line, A, C, threads, pref, wave_A, wave_B, us, p); line rows
    go before the tiling of the same length (the planner takes the first
    entry that applies)."""
    rows = sorted(rows, key=lambda r: (r[0], -r[1]))
    with open(INC, "w") as f:
        f.write("// Generated by gpulucas_amd/tune.py from GPU measurements (MI355X):\n")
        f.write("// {N, A, C, threads, radix preference, column / row wave factor, line kernels}\n")
        for N, line, A, C, T, pref, wa, wb, us, p in rows:
            f.write(f"{{{N}, {A}, {C}, {T}, {pref}, {wa}, {wb}, {line}}},  // {us:.2f} us/iteration at p={p}\n")
    print(f"wrote {len(rows)} entries to {INC}")


def _current_rows():
    rows = []
    with open(INC) as f:
        lines = [l for l in f if l.startswith("{")]
    for (N, A, C, T, pref, wa, wb, line, p), text in zip(tuned_entries(), lines):
        rows.append((N, line, A, C, T, pref, wa, wb, _comment_us(text) or 0.0, p))
    return rows


def emit(paths):
    table = {}
    for path in paths:
        with open(path) as f:
            table.update(json.load(f))
    keep = [r for r in _current_rows() if r[1] == 1]
    rows = [(int(v["N"]), 0, int(v["A"]), int(v["C"]), int(v["threads"]), int(v.get("radix_pref", 8)),
             int(v.get("wave_A", -1)), int(v.get("wave_B", -1)), 1e3 * float(v["ms_per_iter"]), int(v["p"]))
            for v in table.values()]
    _write_inc(rows + keep)


def measure_line(pairs, iters, out, margin=0.02):
    """For each (p, N): the table's plan and every line-kernel plan, timed
    over `iters` graph-launched iterations after their graphs are built."""
    from gpulucas_amd import lucas

    results = {}
    for p, N in pairs:
        cands = lucas.line_plan_list(p, N)
        if not cands:
            continue
        t0 = time.time()

        def timed(ctx):
            ctx.prepare(iters)
            ctx.time_iterations(40)
            return 1e3 * ctx.time_iterations(iters)

        tiled = [r for r in tuned_entries() if r[0] == N and r[7] == 0]
        plan = None
        if tiled:
            _, A, C, T, pref, wa, wb, _, _ = tiled[0]
            plan = {"p": p, "N": N, "A": A, "C": C, "threads": T, "radix_pref": pref, "wave_A": wa, "wave_B": wb}
        with lucas.Context(p, N, plan=plan) as ctx:
            base = timed(ctx)
        best = None
        every = []
        for c in cands:
            # c["line"]: 1 one wave per row, 3 two; +1 non-temporal spectrum loads and stores
            for kind in (c["line"], c["line"] + 1):
                with lucas.Context(p, N, plan={"p": p, "N": N, "A": c["A"], "C": c["C"], "line": kind}) as ctx:
                    us = timed(ctx)
                every.append({"A": c["A"], "C": c["C"], "line": kind, "us": round(us, 3)})
                if best is None or us < best[0]:
                    best = (us, c, kind)
        us, c, kind = best
        results[str(N)] = {"p": p, "N": N, "tiled_us": base, "line_us": us, "A": c["A"], "C": c["C"],
                           "threads": c["threads"], "line": kind, "wins": us < base * (1 - margin),
                           "every": every}
        print(f"N={N} p={p}: tiled {base:.2f} us, line{kind} A={c['A']} K={c['C']} {us:.2f} us "
              f"({'line' if us < base * (1 - margin) else 'tiled'}, {time.time() - t0:.1f}s)", flush=True)
        with open(out, "w") as f:
            json.dump(results, f, indent=1)
    return results


def emit_line(paths):
    """Adds a line entry for every length whose line plan won, replacing the
    table's earlier line entries."""
    table = {}
    for path in paths:
        with open(path) as f:
            table.update(json.load(f))
    rows = [r for r in _current_rows() if r[1] == 0]
    for v in table.values():
        if v["wins"]:
            rows.append((int(v["N"]), int(v.get("line", 1)), int(v["A"]), int(v["C"]), int(v["threads"]), 8, 0, 0,
                         float(v["line_us"]), int(v["p"])))
    _write_inc(rows)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--emit", nargs="*", default=None)
    ap.add_argument("--line", action="store_true", help="time the line-kernel plans against the table's tilings")
    ap.add_argument("--emit-line", nargs="*", default=None)
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
    if a.emit_line is not None:
        emit_line(a.emit_line)
        return
    if a.lengths:
        pairs = [tuple(int(x) for x in s.split(":")) for s in a.lengths.split(",")]
    elif a.range:
        known = {N for _, N in standard_lengths()}
        pairs = [(p, N) for p, N in range_lengths() if N not in known]
    else:
        pairs = standard_lengths(serving=not a.no_serving)
    if a.line:
        if not a.lengths:
            pairs = sorted({(r[8], r[0]) for r in tuned_entries()})
        measure_line(pairs, a.iters, a.out or "line.json")
        return
    measure(pairs, a.iters, a.max_candidates, a.out or "tune.json")


if __name__ == "__main__":
    sys.exit(main())
