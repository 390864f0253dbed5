"""Command-line driver (reference main, gpuLucas.cu:387-639): argument
handling without a GPU."""
import os
import subprocess

from gpulucas_amd import build

CLI = build.CLI


def run(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


def test_help_and_usage_errors():
    r = run("-h")
    assert r.returncode == 0 and "Usage:" in r.stdout and "-c k" in r.stdout
    assert run().returncode == 2          # no exponent
    assert run("-x", "86243").returncode == 2
    assert run("-c", "0", "86243").returncode == 2
    assert os.access(CLI, os.X_OK)


def test_exponent_argument_checks():
    # the reference's bound on the exponent (gpuLucas.cu:437-441)
    for args in (("3000000000",), ("-n", "99999999999"), ("2147483648",)):
        r = run(*args)
        assert r.returncode == 2 and "testPrime too large" in r.stderr, args
    r = run("86243x")
    assert r.returncode == 2 and "Non-option argument 86243x" in r.stderr
    assert run("-n", "abc").returncode == 2
