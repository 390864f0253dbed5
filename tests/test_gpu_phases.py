"""Phase profile: the profiling library's stamps, read through the C ABI.

The default library must refuse (ERR_UNSUPPORTED).  The profiling library
runs in a child process (one library per process): every workgroup of the
last column and row launch reports six ordered stamps.
"""
import json
import os
import subprocess
import sys

import pytest

from gpulucas_amd import lucas

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_default_library_has_no_phase_stamps():
    with lucas.Context(86243, 4608) as ctx:
        with pytest.raises(lucas.LucasError) as e:
            ctx.phase_stamps(2)
        assert e.value.status == lucas.ERR_UNSUPPORTED
        # the context still works afterwards
        ctx.step(10)
        assert ctx.iteration == 10


def test_phase_profile_tool():
    cases = [(86243, 4608), (2976221, 163840)]
    cmd = [sys.executable, "-m", "gpulucas_amd.phases", "--json", "--iters", "8"]
    for p, N in cases:
        cmd += ["--case", f"{p}:{N}"]
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    rows = [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
    assert [(x["p"], x["N"]) for x in rows] == cases
    for x in rows:
        info = lucas.plan(x["p"], x["N"])
        col, row = x["column_kernel"], x["row_kernel"]
        assert col["workgroups"] == info["col_wgs"] and row["workgroups"] == info["row_wgs"]
        for s in (col, row):
            assert 0.0 < s["workgroup_us_mean"] <= s["workgroup_us_max"] <= s["span_us"] < 1000.0
            assert abs(sum(s["phase_us_mean"].values()) - s["workgroup_us_mean"]) < 0.05
            assert all(v >= 0.0 for v in s["phase_us_mean"].values())
