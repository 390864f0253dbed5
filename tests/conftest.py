import glob
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.join(ROOT, "tests")
for _p in (ROOT, TESTS):
    if _p not in sys.path:
        sys.path.insert(0, _p)

FIXTURES = os.path.join(ROOT, "tests", "golden", "gmp_res64.json")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP engine)")


@pytest.fixture(scope="session")
def gmp_fixtures():
    """GMP Res64 fixtures: gmp_res64.json plus any gmp_res64_*.json parts."""
    data = {}
    for path in [FIXTURES] + sorted(glob.glob(FIXTURES.replace(".json", "_*.json"))):
        if os.path.exists(path):
            with open(path) as f:
                for group, entries in json.load(f).items():
                    data.setdefault(group, {}).update(entries)
    return data


def fixture_res64(fx, group, p, k):
    """(res64 int, zero bool) of s_k for exponent p from the GMP fixtures."""
    e = fx[group][str(p)]
    return int(e["res64"][str(k)], 16), bool(e["zero"][str(k)])
