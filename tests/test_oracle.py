"""Pins the CPU oracles (GMP LL and the IBDWT restatement) before they are
trusted as checkers.

Each check is chosen so a plausible mistake fails it: closed-form residues,
known Mersenne primes, exact integer convolution, a value-preserving carry
invariant, extended-precision weights, and planted mistakes in the oracle.
"""
import math

import mpmath
import numpy as np
import pytest

import oracle
from gpulucas_amd import workloads as wl

# s_1..s_5 are below 2^61 and need no reduction for p >= 62.
CLOSED_FORM = [14, 194, 37634, 1416317954, 2005956546822746114]


@pytest.mark.parametrize("p", [89, 107, 127, 521, 1279, 9973])
def test_gmp_closed_form(p):
    got = oracle.gmp_res64(p, [1, 2, 3, 4, 5])
    assert [g for g, _ in got] == CLOSED_FORM
    # s_6 = s_5^2 - 2 (122 bits) for p >= 123
    if p >= 123:
        s6 = CLOSED_FORM[-1] ** 2 - 2
        assert oracle.gmp_residue(p, 6) == s6 % (2 ** p - 1)


@pytest.mark.parametrize("p", [p for p in wl.MERSENNE_PRIME_EXPONENTS if 7 <= p <= 4423])
def test_gmp_known_primes(p):
    (_, z), = oracle.gmp_res64(p, [p - 2])
    assert z


@pytest.mark.parametrize("p", [11, 23, 29, 37, 1009, 2003, 4001])
def test_gmp_known_composites(p):
    (_, z), = oracle.gmp_res64(p, [p - 2])
    assert not z


def test_gmp_residue_matches_python_bigints():
    p = 607
    s = 4
    m = 2 ** p - 1
    for _ in range(40):
        s = (s * s - 2) % m
    assert oracle.gmp_residue(p, 40) == s
    assert oracle.gmp_res64(p, [40])[0][0] == s & (2 ** 64 - 1)


def test_gmp_against_committed_fixtures(gmp_fixtures):
    # recompute the cheap fixture entries (small p, few iterations)
    checked = 0
    for group, entries in gmp_fixtures.items():
        if group.startswith("_"):
            continue
        for ps, e in entries.items():
            p = int(ps)
            ks = [int(k) for k in e["res64"] if p * int(k) <= 3e8]
            if not ks:
                continue
            got = oracle.gmp_res64(p, ks)
            for k, (r, z) in zip(sorted(ks), got):
                assert f"{r:016x}" == e["res64"][str(k)], (group, p, k)
                assert int(z) == e["zero"][str(k)]
                checked += 1
    assert checked > 10


@pytest.mark.parametrize("p,N", [(1279, 64), (2203, 120), (4423, 240), (9941, 560), (19937, 1080),
                                 (3217, 168), (607, 36)])
def test_ibdwt_oracle_matches_gmp(p, N):
    ks = [1, 2, 5, 17, 60]
    want = oracle.gmp_res64(p, ks)
    x = oracle.s0_digits(N)
    done = 0
    for k, (r, z) in zip(ks, want):
        x, err = oracle.ibdwt_iterate(p, N, x, k - done)
        done = k
        assert oracle.ibdwt_res64(p, N, x) == (r, z), (p, N, k)
        assert err < 0.1


@pytest.mark.parametrize("p,N", [(1279, 64), (4423, 240), (9941, 560), (3217, 168)])
def test_ibdwt_oracle_full_run_prime(p, N):
    x, err = oracle.ibdwt_iterate(p, N, oracle.s0_digits(N), p - 2)
    r, z = oracle.ibdwt_res64(p, N, x)
    assert z and r == 0
    assert err < 0.4


def _random_balanced(p, N, seed):
    rng = np.random.default_rng(seed)
    w, _, _, _ = oracle.tables(p, N)
    half = (1 << (w - 1)).astype(np.int64)
    return (rng.integers(0, 2 * half) - half).astype(np.int32)


@pytest.mark.parametrize("p,N", [(1279, 64), (2203, 120), (4423, 240), (2281, 126), (3217, 168), (9941, 560)])
def test_fft_convolution_equals_exact_integer_convolution(p, N):
    for seed in range(3):
        x = _random_balanced(p, N, seed)
        ce = oracle.ibdwt_conv_exact(p, N, x)
        cf, err = oracle.ibdwt_conv_fft(p, N, x)
        assert err < 0.05
        assert np.array_equal(ce, cf)


@pytest.mark.parametrize("p,N", [(1279, 64), (4423, 240), (9941, 560)])
def test_exact_convolution_is_the_square_minus_two(p, N):
    # sum_j c_j 2^{e_j} == X^2 - 2 (mod M_p), checked with big integers
    x = _random_balanced(p, N, 7)
    e = oracle.offsets(p, N)
    m = 2 ** p - 1
    X = sum(int(v) << int(o) for v, o in zip(x, e)) % m
    c = oracle.ibdwt_conv_exact(p, N, x)
    Y = sum(int(v) << int(o) for v, o in zip(c, e)) % m
    assert Y == (X * X - 2) % m


@pytest.mark.parametrize("p,N", [(1279, 64), (4423, 240), (9941, 560), (2281, 126)])
def test_restoration_preserves_value(p, N):
    # sum d_j 2^{e_j} + sum h_j 2^{e_{j+1}} == sum c_j 2^{e_j}  (mod M_p)
    x = _random_balanced(p, N, 3)
    c = oracle.ibdwt_conv_exact(p, N, x)
    d, h = oracle.ibdwt_restore(p, N, c)
    e = oracle.offsets(p, N)
    e_next = np.append(e[1:], p)  # digit N-1 carries into 2^p == 1
    m = 2 ** p - 1
    lhs = (sum(int(v) << int(o) for v, o in zip(d, e)) + sum(int(v) << int(o) for v, o in zip(h, e_next))) % m
    rhs = sum(int(v) << int(o) for v, o in zip(c, e)) % m
    assert lhs == rhs
    w, _, _, _ = oracle.tables(p, N)
    assert np.all(np.abs(d.astype(np.int64)) <= (1 << (w - 1)))
    assert np.abs(h).max() <= 4


@pytest.mark.parametrize("p,N", [(86243, 4608), (1257787, 65536), (2976221, 163840), (3021377, 163840),
                                 (859433, 49152), (42643801, 2359296)])
def test_widths_exact_equal_reference_double_formula(p, N):
    w, m, _, _ = oracle.tables(p, N)
    wd = oracle.widths_double(p, N)
    assert np.array_equal(w, wd)
    assert int(w.sum()) == p
    assert int((w == p // N + 1).sum()) == p % N
    # HI mask bit i <=> digit j-i is wide (computeBitsPerWordVectors)
    hi = (w == p // N + 1).astype(np.uint8)
    for i in range(8):
        assert np.array_equal((m >> i) & 1, np.roll(hi, i))


@pytest.mark.parametrize("p,N", [(86243, 4608), (2976221, 163840), (1257787, 65536)])
def test_weights_against_mpmath(p, N):
    _, _, a, ai = oracle.tables(p, N)
    mpmath.mp.dps = 50
    for j in [0, 1, 2, 3, 17, N // 3, N // 2 + 1, N - 1]:
        exact = mpmath.power(2, mpmath.ceil(mpmath.mpf(j) * p / N) - mpmath.mpf(j) * p / N)
        assert a[j] == float(exact)
        assert abs(ai[j] - float(1 / (exact * N))) <= 2 * math.ulp(ai[j])
    assert a[0] == 1.0


@pytest.mark.parametrize("mutation", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("p,N", [(1279, 64), (4423, 240)])
def test_planted_mistakes_are_caught(mutation, p, N):
    want = oracle.gmp_res64(p, [40])[0]
    x, _ = oracle.ibdwt_iterate(p, N, oracle.s0_digits(N), 40, mutation=mutation)
    assert oracle.ibdwt_res64(p, N, x) != want


def test_carry_depth_rule():
    # setSliceAndDice, gpuLucas.cu:323-343
    assert [oracle.carry_depth(q) for q in (21, 20, 18, 17, 16, 15, 14, 13, 12, 11, 8)] == \
        [-1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6]


def test_product_bits_restatement_pinned_by_bit_length():
    """computeMaxBitVector's reduction (oracle) against Python's exact
    int.bit_length: largest bit length, and largest 64-bit leading-zero
    count (64 for a zero product)."""
    rng = np.random.default_rng(3)
    for trial in range(50):
        n = int(rng.integers(1, 40))
        bits = rng.integers(0, 53, size=n)
        c = [int(rng.integers(0, 1 << int(b))) * (1 if rng.random() < 0.5 else -1) for b in bits]
        if trial % 5 == 0:
            c[0] = 0
        want_b = max(abs(v).bit_length() for v in c)
        want_z = max(64 - abs(v).bit_length() for v in c)
        assert oracle.product_bits(np.array(c, dtype=np.int64)) == (want_b, want_z)
    # a single power of two on either side of a word boundary
    for k in (0, 1, 31, 32, 52):
        assert oracle.product_bits(np.array([-(1 << k)], dtype=np.int64)) == (k + 1, 63 - k)
