"""Deterministic exponent lists for the standard workloads.

Exponent draws use splitmix64 (Steele, Lea & Flood 2014) so the lists are
reproducible on any machine, and primality uses deterministic Miller-Rabin
with bases {2, 3, 5, 7}, which is exact for n < 3,215,031,751.

Workloads (sizes as used by the tests and bench):
  * ``single``   : p=86243 (N=4608), p=1257787 (N=65536), p=2976221 (N=163840)
  * ``near3m``   : 3 primes drawn from seed 3 in [2.9e6, 3.0e6], 20000 iterations
  * ``batch64``  : 64 distinct primes drawn from seed 4 in [1e6, 5e6], 10000 iterations
  * ``serving``  : Poisson arrivals from seed 5, p log-uniform in [1e5, 4e7],
                   K uniform in [5000, 50000]
"""
from __future__ import annotations

import math

MASK64 = (1 << 64) - 1

# Mersenne-prime exponents up to 4e7 (known answers: final residue 0).
MERSENNE_PRIME_EXPONENTS = (
    2, 3, 5, 7, 13, 17, 19, 31, 61, 89, 107, 127, 521, 607, 1279, 2203, 2281,
    3217, 4253, 4423, 9689, 9941, 11213, 19937, 21701, 23209, 44497, 86243,
    110503, 132049, 216091, 756839, 859433, 1257787, 1398269, 2976221, 3021377,
    6972593, 13466917, 20996011, 24036583, 25964951, 30402457, 32582657,
    37156667,
)


class SplitMix64:
    """splitmix64 generator: state += golden gamma, then two xor-shift-multiply rounds."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform_int(self, lo: int, hi: int) -> int:
        """Integer in [lo, hi] as lo + next() mod (hi - lo + 1)."""
        return lo + self.next() % (hi - lo + 1)

    def uniform01(self) -> float:
        """Float in [0, 1) from the top 53 bits."""
        return (self.next() >> 11) * (1.0 / (1 << 53))


def is_prime(n: int) -> bool:
    """Deterministic Miller-Rabin, bases 2, 3, 5, 7 (exact below 3.2e9)."""
    if n < 2:
        return False
    for sp in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % sp == 0:
            return n == sp
    if n >= 3_215_031_751:
        raise ValueError("is_prime: bases {2,3,5,7} are exact only below 3215031751")
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def next_prime(n: int) -> int:
    while not is_prime(n):
        n += 1
    return n


def near3m_primes(seed: int = 3, count: int = 3, lo: int = 2_900_000, hi: int = 3_000_000):
    rng = SplitMix64(seed)
    out = []
    while len(out) < count:
        p = next_prime(rng.uniform_int(lo, hi))
        if p <= hi and p not in out:
            out.append(p)
    return out


def batch_primes(seed: int = 4, count: int = 64, lo: int = 1_000_000, hi: int = 5_000_000):
    return near3m_primes(seed=seed, count=count, lo=lo, hi=hi)


def serving_jobs(seed: int = 5, count: int = 64, mean_interarrival_s: float = 0.05,
                 p_lo: float = 1e5, p_hi: float = 4e7, k_lo: int = 5000, k_hi: int = 50000):
    """Poisson arrivals: list of (arrival_time_s, p, K), sorted by arrival."""
    rng = SplitMix64(seed)
    t = 0.0
    jobs = []
    lo, hi = math.log(p_lo), math.log(p_hi)
    for _ in range(count):
        u = rng.uniform01()
        t += -math.log(1.0 - u) * mean_interarrival_s
        p = next_prime(int(math.exp(lo + (hi - lo) * rng.uniform01())))
        k = rng.uniform_int(k_lo, k_hi)
        jobs.append((t, p, k))
    return jobs
