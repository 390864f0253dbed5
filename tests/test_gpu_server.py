"""Serving: continuous batching of (p, iterations) requests with HBM-footprint
admission.  Each result must be identical alone and in the mix, and must
equal GMP."""
import os
import threading
import time

import pytest

from conftest import fixture_res64
from gpulucas_amd import lucas, workloads as wl

pytestmark = pytest.mark.gpu

LIMIT = lucas.DEFAULT_ERROR_LIMIT
SMALL_PRIMES = [9689, 9941, 11213, 19937, 21701, 23209, 44497, 86243]
SMALL_COMPOSITES = [9973, 10007, 20011, 30011, 44501, 50021]


def test_server_full_runs_match_gmp(gmp_fixtures):
    with lucas.Server(max_jobs=8, chunk_iters=128) as s:
        ids = {}
        for p in SMALL_PRIMES + SMALL_COMPOSITES:
            ids[s.submit(p, p - 2)] = p
        s.drain(timeout_s=300)
        for jid, p in ids.items():
            r = s.poll(jid)
            assert r["done"] == 1 and r["status"] == lucas.OK
            if p in SMALL_PRIMES:
                assert r["is_zero"] == 1 and r["res64"] == 0, p
            else:
                want, _ = fixture_res64(gmp_fixtures, "small_composite_full", p, p - 2)
                assert r["res64"] == want, p
            assert r["max_roundoff"] < LIMIT
        st = s.stats()
        assert st["peak_jobs"] <= 8 and st["iterations"] >= sum(p - 2 for p in ids.values())


def test_busy_server_batch_plans_match_gmp(gmp_fixtures):
    """More than 16 jobs outstanding: later submissions take the batch plan
    (wider tiles); every result still equals GMP."""
    primes = [p for p in wl.MERSENNE_PRIME_EXPONENTS if 1279 <= p <= 132049]
    jobs = primes + SMALL_COMPOSITES
    assert len(jobs) > 16
    with lucas.Server(chunk_iters=256) as s:
        ids = {s.submit(p, p - 2): p for p in jobs}
        s.drain(timeout_s=300)
        for jid, p in ids.items():
            r = s.poll(jid)
            assert r["done"] == 1 and r["status"] == lucas.OK, p
            if p in primes:
                assert r["is_zero"] == 1 and r["res64"] == 0, p
            else:
                want, _ = fixture_res64(gmp_fixtures, "small_composite_full", p, p - 2)
                assert r["res64"] == want, p
            assert r["max_roundoff"] < LIMIT


def test_server_admission_by_footprint():
    ps = [86243, 110503, 132049, 216091]
    sizes = sorted(lucas.plan(p)["device_bytes"] for p in ps)
    budget = int(2.2 * sizes[-1])
    # most jobs that can ever be resident together under the budget
    fit = max(k for k in range(1, len(ps) + 1) if sum(sizes[:k]) <= budget)
    assert fit < len(ps)
    with lucas.Server(hbm_budget_bytes=budget, chunk_iters=256) as s:
        ids = [s.submit(p, 3000) for p in ps]
        s.drain(timeout_s=300)
        st = s.stats()
        assert 1 <= st["peak_jobs"] <= fit
        res = [s.poll(i) for i in ids]
    for p, r in zip(ps, res):
        with lucas.Context(p, 0) as c:
            c.step(3000)
            assert (r["res64"], bool(r["is_zero"])) == c.res64()


def test_server_poisson_arrivals_alone_equals_mixed():
    # jobs arrive while others are running; each result must match its solo run
    jobs = [(t * 0.2, p, k // 10) for t, p, k in wl.serving_jobs(seed=5, count=12, p_hi=3e6)]
    out = {}
    with lucas.Server(chunk_iters=200) as s:
        t0 = time.time()

        def feed():
            for t, p, k in jobs:
                dt = t - (time.time() - t0)
                if dt > 0:
                    time.sleep(dt)
                out[s.submit(p, k)] = (p, k)

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        s.drain(timeout_s=600)
        results = {jid: s.poll(jid) for jid in out}
        assert s.stats()["peak_jobs"] >= 2  # work overlapped on the device
    lat = sorted(r["finish_s"] - r["submit_s"] for r in results.values())
    print(f"latency p50={lat[len(lat) // 2]:.3f}s max={lat[-1]:.3f}s", flush=True)
    for jid, (p, k) in out.items():
        r = results[jid]
        assert r["status"] == lucas.OK
        with lucas.Context(p, 0) as c:
            c.step(k)
            assert (r["res64"], bool(r["is_zero"])) == c.res64(), (p, k)


def test_server_checkpoints(tmp_path):
    with lucas.Server(chunk_iters=500, checkpoint_dir=str(tmp_path), checkpoint_every=1000) as s:
        jid = s.submit(86243, 5000)
        s.drain(timeout_s=120)
        r = s.poll(jid)
    path = tmp_path / f"86243_{jid}.chk"
    assert path.exists()
    with lucas.Context.load(str(path)) as c:
        assert c.iteration % 1000 == 0 and 0 < c.iteration < 5000
        c.step(5000 - c.iteration)
        assert c.res64() == (r["res64"], bool(r["is_zero"]))


def test_server_checkpoints_of_one_exponent_twice(tmp_path):
    """Two resident jobs of the same exponent keep separate checkpoint files
    (<p>_<id>.chk), each resuming to its own job's result."""
    with lucas.Server(chunk_iters=500, checkpoint_dir=str(tmp_path), checkpoint_every=1000) as s:
        a = s.submit(86243, 3000)
        b = s.submit(86243, 5000)
        s.drain(timeout_s=120)
        ra, rb = s.poll(a), s.poll(b)
    for jid, r, total in ((a, ra, 3000), (b, rb, 5000)):
        path = tmp_path / f"86243_{jid}.chk"
        assert path.exists()
        with lucas.Context.load(str(path)) as c:
            assert 0 < c.iteration < total
            c.step(total - c.iteration)
            assert c.res64() == (r["res64"], bool(r["is_zero"]))


def test_server_reports_roundoff_breach_per_job():
    """A served job whose roundoff reaches the server's limit finishes with
    ERR_ROUNDOFF (its residue is not trusted); the gate is per job."""
    with lucas.Server(chunk_iters=256, error_limit=1e-9) as s:
        jid = s.submit(86243, 1000)
        s.drain(timeout_s=120)
        r = s.poll(jid)
    assert r["done"] == 1 and r["status"] == lucas.ERR_ROUNDOFF and r["max_roundoff"] > 0


def test_config5_serving_full_size_vs_gmp(gmp_fixtures):
    """The serving workload at its stated size: the first 48 Poisson jobs of
    seed 5 (p log-uniform in [1e5, 4e7], N from 6K to 2.2M words, K uniform
    in [5000, 50000], each at its full K), arriving in order over 20 s while
    earlier jobs run.  Every served residue equals GMP's; roundoff stays
    under the gate on every job; work overlaps on the device."""
    jobs = wl.serving_jobs(seed=5, count=48)
    fx = gmp_fixtures.get("c5_serving", {})
    missing = [p for _, p, _ in jobs if str(p) not in fx]
    assert not missing, f"GMP fixtures missing for {missing}"
    scale = 20.0 / jobs[-1][0]
    out = {}
    with lucas.Server(chunk_iters=256) as s:
        t0 = time.time()

        def feed():
            for t, p, k in jobs:
                dt = t * scale - (time.time() - t0)
                if dt > 0:
                    time.sleep(dt)
                out[s.submit(p, k)] = (p, k)

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        s.drain(timeout_s=900)
        results = {jid: s.poll(jid) for jid in out}
        st = s.stats()
    lat = sorted(r["finish_s"] - r["submit_s"] for r in results.values())
    print(f"48 jobs: p50={lat[24]:.2f}s max={lat[-1]:.2f}s peak_jobs={st['peak_jobs']}", flush=True)
    assert st["peak_jobs"] >= 2
    for jid, (p, k) in out.items():
        r = results[jid]
        assert r["done"] == 1 and r["status"] == lucas.OK, (p, k, r["status"])
        assert r["max_roundoff"] < LIMIT, (p, r["max_roundoff"])
        assert (r["res64"], bool(r["is_zero"])) == fixture_res64(gmp_fixtures, "c5_serving", p, k), (p, k)
