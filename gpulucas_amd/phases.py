"""Phase profile of the two iteration kernels, from the profiling library.

    python -m gpulucas_amd.phases --case 2976221:163840 --case 37156667:2097152
    python -m gpulucas_amd.phases --case 2976221:163840:256:2   # line kernels, A = 256, K = 2
    python -m gpulucas_amd.phases --case 37156667:2097152:512:4:4  # ... line kind 4 (split rows, non-temporal)

Loads libgpulucas_phases.so (the kernels with their phase stamps compiled
in), runs `--iters` iterations of each exponent at its planned shape and
prints, per kernel, the spread of workgroup start times, the workgroup
duration and the mean and maximum time of each phase over the workgroups of
the last launch, one JSON line per case (`--json`) or a table.  Stamps are
the GPU's 100 MHz real-time counter (10 ns ticks), read by work-item 0 after
the barrier that closes each phase: a phase is the time the whole workgroup
spent in it.  The last phase ("store") only covers issuing the stores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

TICK_US = 0.01  # 100 MHz real-time counter
COL_PHASES = ["load", "inverse FFT + unweight", "carry", "forward FFT", "store"]
ROW_PHASES = ["load", "forward FFT", "square", "inverse FFT", "store"]


def summarize(stamps: np.ndarray, names: list[str]) -> dict:
    """Span, start spread, workgroup duration and per-phase mean/max (us)."""
    if stamps.shape[0] == 0 or stamps.shape[1] != len(names) + 1:
        raise ValueError(f"expected (workgroups, {len(names) + 1}) stamps, got {stamps.shape}")
    d = np.diff(stamps, axis=1)
    if (d < 0).any():
        raise ValueError("phase stamps out of order")
    t0 = stamps[:, 0].min()
    starts = (stamps[:, 0] - t0) * TICK_US
    total = (stamps[:, -1] - stamps[:, 0]) * TICK_US
    ph = d * TICK_US
    return {
        "workgroups": int(stamps.shape[0]),
        "span_us": round(float((stamps[:, -1].max() - t0) * TICK_US), 2),
        "start_spread_us": round(float(starts.max()), 2),
        "workgroup_us_mean": round(float(total.mean()), 2),
        "workgroup_us_max": round(float(total.max()), 2),
        "phase_us_mean": {n: round(float(ph[:, k].mean()), 2) for k, n in enumerate(names)},
        "phase_us_max": {n: round(float(ph[:, k].max()), 2) for k, n in enumerate(names)},
    }


def profile(p: int, N: int, iters: int = 20, line: tuple[int, ...] | None = None) -> dict:
    """Per-phase times at the planned shape, or with `line` = (A, K[, kind])
    at the line kernels' shape with A rows and K own columns per column
    workgroup (kind as lucas_plan_info.line, default 1)."""
    from gpulucas_amd import lucas

    plan = {"p": p, "N": N, "A": line[0], "C": line[1], "line": line[2] if len(line) > 2 else 1} if line else None
    with lucas.Context(p, N, plan=plan) as ctx:
        info = ctx.info
        ctx.step(64)
        col, row = ctx.phase_stamps(iters)
    out = {"p": p, "N": N}
    out.update({k: info[k] for k in ("A", "B", "C", "halo", "threads", "wave_A", "wave_B", "col_dense", "line")})
    out["column_kernel"] = summarize(col, COL_PHASES)
    out["row_kernel"] = summarize(row, ROW_PHASES)
    return out


def _table(r: dict) -> str:
    lines = [f"p={r['p']} N={r['N']} A={r['A']} B={r['B']} C={r['C']} H={r['halo']} T={r['threads']} "
             f"W={r['wave_A']}/{r['wave_B']} dense={r['col_dense']} line={r['line']}"]
    for kern in ("column_kernel", "row_kernel"):
        s = r[kern]
        lines.append(f"  {kern}: {s['workgroups']} workgroups, span {s['span_us']} us, starts within "
                     f"{s['start_spread_us']} us, workgroup {s['workgroup_us_mean']} us mean / "
                     f"{s['workgroup_us_max']} us max")
        for n in s["phase_us_mean"]:
            lines.append(f"    {n:24s} {s['phase_us_mean'][n]:7.2f} mean {s['phase_us_max'][n]:7.2f} max")
    return "\n".join(lines)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--case", action="append", default=[],
                    help="p:N, or p:N:A:K[:kind] for the line kernels (repeatable)")
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--json", action="store_true")
    a = ap.parse_args(argv)
    from gpulucas_amd import build, lucas

    if lucas._LIB is not None and lucas.LIB_PATH != build.PHASES_LIB:
        raise RuntimeError("the default library is already loaded in this process")
    if not os.path.exists(build.PHASES_LIB):
        raise ImportError(f"{build.PHASES_LIB} is not built; run `python -m gpulucas_amd.build`")
    lucas.LIB_PATH = build.PHASES_LIB
    cases = a.case or ["2976221:163840"]
    for c in cases:
        f = [int(x) for x in c.split(":")]
        p, N = f[0], f[1]
        line = tuple(f[2:]) if len(f) >= 4 else None
        print(f"case p={p} N={N}" + (f" line {line}" if line else ""), file=sys.stderr, flush=True)
        r = profile(p, N, a.iters, line)
        print(json.dumps(r) if a.json else _table(r), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
