"""Batched exponents: 64 primes of uneven N advanced by one launch per kernel
per iteration; every residue must equal GMP and each job's digits must be
the same alone and in the mix."""
import numpy as np
import pytest

from conftest import fixture_res64
from gpulucas_amd import lucas, workloads as wl

pytestmark = pytest.mark.gpu


def test_batch64_residues_vs_gmp(gmp_fixtures):
    ps = wl.batch_primes()
    with lucas.Batch(ps) as b:
        assert len(set(b.Ns)) > 8  # genuinely mixed lengths
        for k in (1000, 10000):
            b.step(k - (1000 if k == 10000 else 0))
            res, zero, err = b.results()
            for p, r, z, e in zip(ps, res, zero, err):
                want, wz = fixture_res64(gmp_fixtures, "c4_batch64", p, k)
                assert (r, z) == (want, wz), (p, k)
                assert e < lucas.DEFAULT_ERROR_LIMIT, (p, e)
            print(f"k={k} max roundoff {max(err):.4f}", flush=True)


def test_batch_alone_equals_mixed():
    ps = [86243, 110503, 132049, 216091, 21701, 44497]
    Ns = [0, 0, 0, 0, 0, 0]
    with lucas.Batch(ps, Ns) as b:
        b.step(300)
        mixed = [b.digits(i) for i in range(len(ps))]
    for i, p in enumerate(ps):
        with lucas.Context(p, 0) as c:
            c.step(300)
            assert np.array_equal(c.digits(), mixed[i]), p


def test_batch_full_runs_of_primes():
    ps = [9689, 9941, 11213, 19937, 21701, 23209, 44497]
    iters = [p - 2 for p in ps]
    # one batch step count for all: run the smallest to completion first
    with lucas.Batch(ps) as b:
        done = 0
        for target in sorted(set(iters)):
            b.step(target - done)
            done = target
            res, zero, err = b.results()
            for p, r, z in zip(ps, res, zero):
                if p - 2 == target:
                    assert z and r == 0, p
