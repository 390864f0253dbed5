"""Command-line driver on the GPU: result lines, results.txt, checkpoint on
SIGINT and resume (gpuLucas.cu:365-369, 598-606, 1414-1429, 1514-1547)."""
import os
import signal
import subprocess
import time

import pytest

from conftest import fixture_res64
from gpulucas_amd import build

pytestmark = pytest.mark.gpu
CLI = build.CLI


def test_cli_prime_and_composite(tmp_path, gmp_fixtures):
    r = subprocess.run([CLI, "-f", "4608", "-c", "20000", "86243"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    assert "M( 86243 )P, n = 4608" in r.stdout
    want, _ = fixture_res64(gmp_fixtures, "small_composite_full", 9973, 9971)
    r = subprocess.run([CLI, "-q", "9973"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert f"M( 9973 )C, 0x{want:016x}" in r.stdout
    lines = (tmp_path / "results.txt").read_text().splitlines()
    assert lines[0].startswith("M( 86243 )P") and lines[1].startswith("M( 9973 )C")
    assert not (tmp_path / "86243.chk").exists()  # removed after a finished run


def test_cli_interrupt_then_resume(tmp_path):
    args = [CLI, "-f", "49152", "-c", "20000", "859433"]
    proc = subprocess.Popen(args, cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    ck = tmp_path / "859433.chk"
    deadline = time.time() + 120
    while not ck.exists() and time.time() < deadline and proc.poll() is None:
        time.sleep(0.2)
    proc.send_signal(signal.SIGINT)
    out, err = proc.communicate(timeout=120)
    assert proc.returncode == 3, (out, err)
    assert "Interrupted at iteration" in out and ck.exists()
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Resuming M859433" in r.stdout
    assert "M( 859433 )P, n = 49152" in r.stdout


def test_cli_refuses_a_length_past_the_roundoff_gate(tmp_path):
    """-f with a length far too short for the exponent (20 bits per word at
    N = 2^20): the error trial trips the gate, nothing is written to
    results.txt and the exit status is non-zero."""
    from gpulucas_amd import workloads as wl
    N = 1 << 20
    p = wl.next_prime(20 * N)
    r = subprocess.run([CLI, "-q", "-f", str(N), str(p)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "roundoff" in r.stderr
    assert not (tmp_path / "results.txt").exists()
