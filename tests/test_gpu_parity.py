"""GPU parity: the HIP engine (through the C ABI) against the CPU oracles.

* digit vectors after k iterations are bit-identical to the oracle's
  restatement of the reference (same carry-save restoration, same ripple);
* Res64 equals GMP at the iterations the workloads name; Mersenne-prime
  exponents end at residue 0;
* max roundoff stays below the library's gate (0.30, the reference's
  ERROR_LIMIT) on every case, and the gate fires where it must.
"""
import numpy as np
import pytest

import oracle
from conftest import fixture_res64
from gpulucas_amd import lucas

pytestmark = pytest.mark.gpu

ROUNDOFF_LIMIT = lucas.DEFAULT_ERROR_LIMIT


def random_balanced(p, N, seed):
    rng = np.random.default_rng(seed)
    w, _, _ = lucas.digit_tables(p, N)
    half = (1 << (w - 1)).astype(np.int64)
    return (rng.integers(0, 2 * half) - half).astype(np.int32)


# (p, N, A, C): radix 2/3/5/7 mixes, odd A, C=1..16, halo wider than C
SHAPES = [
    (1279, 64, 0, 0), (1279, 64, 4, 1), (1279, 64, 8, 4),
    (2203, 120, 0, 0), (4423, 240, 0, 0), (4423, 240, 15, 8),
    (9941, 560, 0, 0), (9941, 560, 7, 4), (19937, 1080, 0, 0),
    (21701, 1176, 0, 0), (44497, 2400, 0, 0), (86243, 4608, 0, 0),
    (86243, 4608, 48, 16), (86243, 4608, 72, 8), (110503, 6048, 0, 0),
]


@pytest.mark.parametrize("p,N,A,C", SHAPES)
def test_digits_bit_exact_vs_oracle_from_s0(p, N, A, C):
    print(f"case p={p} N={N} A={A} C={C}", flush=True)
    with lucas.Context(p, N, A=A, C_=C) as ctx:
        x = oracle.s0_digits(N)
        assert np.array_equal(ctx.digits(), x)
        done = 0
        for k in (1, 2, 3, 7, 12):
            ctx.step(k - done)
            x, _ = oracle.ibdwt_iterate(p, N, x, k - done)
            done = k
            assert np.array_equal(ctx.digits(), x), f"digits differ after {k} iterations"
        assert ctx.max_roundoff() < ROUNDOFF_LIMIT


# (p, N, A, C, threads): wave-register factors P = 1, 2, 4 after 0..2 Stockham
# passes, and a dense (two workgroups per CU) column kernel
WAVE_SHAPES = [
    (294919, 16384, 64, 4, 512),     # columns 64 (P=1), rows 128 (P=2)
    (218737, 12288, 192, 2, 1024),   # columns 3 x 64
    (294919, 16384, 512, 2, 512),    # columns 2 x 256 (P=4)
    (368647, 20480, 16, 8, 1024),    # rows 5 x 128
    (2424833, 131072, 64, 2, 1024),  # 512 column workgroups: dense column kernel
]


@pytest.mark.parametrize("p,N,A,C,T", WAVE_SHAPES)
def test_wave_factor_shapes_bit_exact_vs_oracle(p, N, A, C, T):
    print(f"case p={p} N={N} A={A} C={C} T={T}", flush=True)
    plan = [x for x in lucas.plan_list(p, N) if x["A"] == A and x["C"] == C and x["threads"] == T][0]
    assert plan["wave_A"] or plan["wave_B"]
    if N == 131072:
        assert plan["col_dense"] == 1
    x = random_balanced(p, N, 12)
    with lucas.Context(p, N, plan=plan) as ctx:
        ctx.set_digits(x, 0)
        ctx.step(2)
        y, _ = oracle.ibdwt_iterate(p, N, x, 2)
        assert np.array_equal(ctx.digits(), y)
        assert ctx.max_roundoff() < ROUNDOFF_LIMIT


def tiled_plan(p, N):
    """The table's tiling of the general kernels for length N (the plan a
    batch uses), as a plan for exponent p."""
    from gpulucas_amd import tune
    N_, A, C, T, pref, wA, wB, line, _ = [r for r in tune.tuned_entries() if r[0] == N and r[7] == 0][0]
    return {"p": p, "N": N, "A": A, "C": C, "threads": T, "radix_pref": pref, "wave_A": wA, "wave_B": wB,
            "col_dense": 0, "line": 0}


def _tuned_pairs(max_n):
    from gpulucas_amd import tune
    return sorted({(row[8], row[0]) for row in tune.tuned_entries() if row[0] <= max_n})


@pytest.mark.parametrize("p,N", _tuned_pairs(294912))
def test_tuned_shapes_bit_exact_vs_oracle(p, N):
    """Every measured shape of the planner's table up to N = 294912 (the
    batch and serving lengths, with every wave-factor choice the tuner made)
    matches the oracle's digits from random steady-state digits."""
    info = lucas.plan(p, N)
    print(f"case p={p} N={N} A={info['A']} C={info['C']} T={info['threads']} "
          f"W={info['wave_A']}/{info['wave_B']}", flush=True)
    x = random_balanced(p, N, 13)
    with lucas.Context(p, N) as ctx:
        ctx.set_digits(x, 0)
        ctx.step(1)
        y, _ = oracle.ibdwt_iterate(p, N, x, 1)
        assert np.array_equal(ctx.digits(), y)
        assert 0.0 < ctx.max_roundoff() < ROUNDOFF_LIMIT


@pytest.mark.parametrize("p,N,A,C", SHAPES[::2])
def test_digits_bit_exact_vs_oracle_steady_state(p, N, A, C):
    print(f"case p={p} N={N} A={A} C={C}", flush=True)
    x = random_balanced(p, N, 11)
    with lucas.Context(p, N, A=A, C_=C) as ctx:
        ctx.set_digits(x, 0)
        ctx.step(3)
        y, _ = oracle.ibdwt_iterate(p, N, x, 3)
        assert np.array_equal(ctx.digits(), y)
        # random full-size digits always leave some FP64 rounding error: a
        # zero would mean the device-side maximum lost its inputs
        assert 0.0 < ctx.max_roundoff() < ROUNDOFF_LIMIT


@pytest.mark.parametrize("p,N", [(1279, 64), (2203, 120), (4423, 240), (9941, 560), (19937, 1080),
                                 (44497, 2400), (86243, 4608), (110503, 6048), (132049, 7168)])
def test_full_run_mersenne_primes(p, N):
    print(f"case p={p} N={N}", flush=True)
    r = lucas.mersenneTest(p, N)
    assert r["status"] == lucas.OK
    assert r["is_prime"] == 1 and r["res64"] == 0
    assert r["iterations"] == p - 2
    assert r["max_roundoff"] < ROUNDOFF_LIMIT


@pytest.mark.parametrize("p", [61, 89, 107, 127, 521, 607])
def test_smallest_mersenne_primes(p):
    """The bottom of the range (N = 16 .. 30, carry depth up to 6): the
    Mersenne primes end at residue 0."""
    r = lucas.mersenneTest(p, 0)
    assert r["status"] == lucas.OK and r["is_prime"] == 1 and r["res64"] == 0
    assert r["iterations"] == p - 2


@pytest.mark.parametrize("p", [37, 41, 43, 47, 53, 59, 67, 71, 101, 509])
def test_smallest_composites_vs_gmp(p):
    """The smallest exponents with a tiling: full runs, Res64 against GMP."""
    r = lucas.mersenneTest(p, 0)
    want, zero = oracle.gmp_res64(p, [p - 2])[0]
    assert (r["res64"], bool(r["is_prime"])) == (want, zero) and not zero
    assert r["max_roundoff"] < ROUNDOFF_LIMIT


@pytest.mark.parametrize("p", [1009, 2003, 4001, 9973, 10007, 20011, 30011, 44501, 50021])
def test_full_run_composites_vs_gmp(p, gmp_fixtures):
    print(f"case p={p}", flush=True)
    r = lucas.mersenneTest(p, 0)
    want, zero = fixture_res64(gmp_fixtures, "small_composite_full", p, p - 2)
    assert r["res64"] == want and not zero and r["is_prime"] == 0
    assert r["max_roundoff"] < ROUNDOFF_LIMIT


def test_config1_p86243_n4608(gmp_fixtures):
    p, N = 86243, 4608
    with lucas.Context(p, N) as ctx:
        for k in (1000, 10000, 86241):
            ctx.step(k - ctx.iteration)
            want, zero = fixture_res64(gmp_fixtures, "c1_p86243", p, k)
            assert ctx.res64() == (want, zero), k
        assert ctx.res64() == (0, True)
        assert ctx.max_roundoff() < ROUNDOFF_LIMIT


def test_step_sizes_across_replay_graph_boundaries(gmp_fixtures):
    """One step of 1000 iterations (graphs of 128 pairs, then the binary
    digits of the rest) and the same 1000 split at every graph-size boundary
    (whole-step graphs up to 64 iterations, 32-pair graphs below 512) reach
    the same digits and GMP's residue."""
    p, N = 86243, 4608
    want, zero = fixture_res64(gmp_fixtures, "c1_p86243", p, 1000)
    with lucas.Context(p, N) as a, lucas.Context(p, N) as b:
        a.step(1000)
        for k in (1, 31, 32, 33, 63, 64, 65, 127, 128, 129, 327):
            b.step(k)
        assert a.iteration == b.iteration == 1000
        assert np.array_equal(a.digits(), b.digits())
        assert a.res64() == (want, zero)


def test_config2_p1257787_n65536_full(gmp_fixtures):
    p, N = 1257787, 65536
    with lucas.Context(p, N) as ctx:
        ks = [1000, 10000] + [100000 * m for m in range(1, 13)] + [1257785]
        for k in ks:
            ctx.step(k - ctx.iteration)
            want, zero = fixture_res64(gmp_fixtures, "c2_p1257787", p, k)
            assert ctx.res64() == (want, zero), k
            print(f"k={k} roundoff={ctx.max_roundoff():.4f}", flush=True)
        assert ctx.res64() == (0, True)
        assert ctx.max_roundoff() < ROUNDOFF_LIMIT


def test_config3_p2976221_n163840(gmp_fixtures):
    p, N = 2976221, 163840
    with lucas.Context(p, N) as ctx:
        for k in (1000, 5000, 20000, 100000):
            ctx.step(k - ctx.iteration)
            want, zero = fixture_res64(gmp_fixtures, "c3_p2976221", p, k)
            assert ctx.res64() == (want, zero), k
        assert ctx.max_roundoff() < ROUNDOFF_LIMIT


def test_config3_p2976221_full_run_is_prime(gmp_fixtures):
    """All 2,976,219 iterations of M2976221 at N=163840: residues at the GMP
    checkpoints on the way, final residue 0."""
    p, N = 2976221, 163840
    with lucas.Context(p, N) as ctx:
        for k in (20000, 100000):
            ctx.step(k - ctx.iteration)
            want, zero = fixture_res64(gmp_fixtures, "c3_p2976221", p, k)
            assert ctx.res64() == (want, zero), k
        while ctx.iteration < p - 2:
            ctx.step(min(500000, p - 2 - ctx.iteration))
            print(f"k={ctx.iteration} roundoff={ctx.max_roundoff():.4f}", flush=True)
        assert ctx.res64() == (0, True)
        assert ctx.max_roundoff() < ROUNDOFF_LIMIT


@pytest.mark.parametrize("p,k", [(9600037, 1000), (9600037, 5000), (20996011, 100), (20996011, 2000),
                                 (37156667, 100), (37156667, 2000), (37732501, 2000)])
def test_long_transforms_vs_gmp(p, k, gmp_fixtures):
    """N = 2^19 .. 2293760 (the top of the serving range, carry depth 3 at
    q = 17): Res64 against GMP."""
    N = lucas.default_N(p)
    print(f"case p={p} N={N} k={k}", flush=True)
    with lucas.Context(p, N) as ctx:
        ctx.step(k)
        want, zero = fixture_res64(gmp_fixtures, "large", p, k)
        assert ctx.res64() == (want, zero)
        assert ctx.max_roundoff() < ROUNDOFF_LIMIT


@pytest.mark.parametrize("p,N,k", [(20996011, 1572864, 100), (20996011, 1572864, 2000),
                                   (37156667, 2097152, 100), (37156667, 2097152, 2000)])
def test_dense_column_kernel_vs_gmp(p, N, k, gmp_fixtures):
    """Lengths whose column grid has more 1024-item tiles than CUs run the
    dense (64-VGPR, two workgroups per CU) column kernel in their tiled plan
    (the one batches use): Res64 against GMP."""
    print(f"case p={p} N={N} k={k}", flush=True)
    with lucas.Context(p, N, plan=tiled_plan(p, N)) as ctx:
        info = ctx.info
        assert info["col_dense"] == 1 and info["threads"] == 1024 and info["col_wgs"] > 256
        ctx.step(k)
        want, zero = fixture_res64(gmp_fixtures, "large", p, k)
        assert ctx.res64() == (want, zero)
        assert ctx.max_roundoff() < ROUNDOFF_LIMIT


@pytest.mark.parametrize("p,N", [(57885161, 3670016), (100000007, 6291456), (136279841, 8064000)])
def test_longest_lengths_vs_gmp(p, N, gmp_fixtures):
    """The longest lengths: M48 at its tuned shape and, past the ordinary
    tilings (wide tiles, up to 8 words per work-item in radix 3/5/7/8
    stages), up to M52 = 2^136279841 - 1: Res64 after 100 iterations
    against GMP."""
    assert lucas.default_N(p) == N
    info = lucas.plan(p, N)
    print(f"case p={p} N={N} A={info['A']} B={info['B']} C={info['C']} T={info['threads']}", flush=True)
    with lucas.Context(p, N) as ctx:
        ctx.step(100)
        want, zero = fixture_res64(gmp_fixtures, "longest", p, 100)
        assert ctx.res64() == (want, zero)
        assert ctx.max_roundoff() < ROUNDOFF_LIMIT


def test_config3_near3m_primes(gmp_fixtures):
    from gpulucas_amd import workloads as wl
    for p in wl.near3m_primes():
        print(f"case p={p}", flush=True)
        with lucas.Context(p, 163840) as ctx:
            ctx.step(20000)
            want, zero = fixture_res64(gmp_fixtures, "c3_near3m", p, 20000)
            assert ctx.res64() == (want, zero)
            assert ctx.max_roundoff() < ROUNDOFF_LIMIT


def test_error_test_contract():
    ms = lucas.errorTest(86243, 4608, 200)
    assert ms > 0


def test_roundoff_gate_fires_on_an_explicit_limit():
    """errorTrial's contract (gpuLucas.cu:904-907, 1239-1242): a limit below
    the measured roundoff returns ERR_ROUNDOFF and errorTest returns -1.0;
    mersenneTest stops at the first chunk and reports the breach."""
    ok = lucas.error_test(86243, 4608, 200)
    assert ok["status"] == lucas.OK and 0.0 < ok["max_roundoff"] < ROUNDOFF_LIMIT
    bad = lucas.error_test(86243, 4608, 200, error_limit=1e-9)
    assert bad["status"] == lucas.ERR_ROUNDOFF and bad["max_roundoff"] == ok["max_roundoff"]
    assert lucas.errorTest(86243, 4608, 200, error_limit=1e-9) == -1.0
    r = lucas.mersenne_test(86243, 4608, error_limit=1e-9)
    assert r["status"] == lucas.ERR_ROUNDOFF and r["is_prime"] == 0
    assert 0 < r["iterations"] < 86241  # stopped at the first gate check


def test_roundoff_gate_fires_past_the_calibrated_line():
    """A length far too short for its exponent: p = 20.0 bits per word at
    N = 2^20, 1.8 bits above the calibrated line.  The roundoff must reach
    the 0.30 gate with no help from the caller."""
    from gpulucas_amd import workloads as wl
    N = 1 << 20
    p = wl.next_prime(20 * N)
    info = lucas.plan(p, N)
    assert info["q"] == 20 and p / N - 18.2 > 1.5
    r = lucas.error_test(p, N, 200)
    print(f"p={p} N={N} roundoff={r['max_roundoff']:.4f}", flush=True)
    assert r["status"] == lucas.ERR_ROUNDOFF and r["max_roundoff"] >= ROUNDOFF_LIMIT
    assert lucas.errorTest(p, N, 200) == -1.0


def test_product_bits_match_exact_convolution():
    """computeMaxBitVector's diagnostic (gpuLucas.cu:826-836): after one
    iteration from random digits, the device maximum of bit lengths and of
    64-bit leading-zero counts of the rounded products equals the exact
    weighted convolution's."""
    for p, N in [(4423, 240), (9941, 560), (19937, 1080)]:
        x = random_balanced(p, N, 31)
        want = oracle.product_bits(oracle.ibdwt_conv_exact(p, N, x))
        with lucas.Context(p, N) as ctx:
            ctx.set_digits(x, 0)
            ctx.product_bits(reset=True)
            ctx.step(1)
            assert ctx.product_bits() == want, (p, N)


@pytest.mark.parametrize("p,N", [(86243, 4608), (1257787, 65536), (2976221, 163840)])
def test_product_bits_at_config_shapes(p, N):
    """The same at the configuration lengths against the oracle's
    long-double convolution (exact integers at these roundoffs)."""
    x = random_balanced(p, N, 32)
    c, err = oracle.ibdwt_conv_fft(p, N, x)
    assert err < 0.25
    want = oracle.product_bits(c)
    with lucas.Context(p, N) as ctx:
        ctx.set_digits(x, 0)
        ctx.product_bits(reset=True)
        ctx.step(1)
        got = ctx.product_bits()
    print(f"p={p} N={N} product bits {got}", flush=True)
    assert got == want
    assert got[0] <= 53


@pytest.mark.parametrize("p,group,k", [(2976221, "c3_p2976221", 20000), (37156667, "large", 2000)])
def test_autotune_n_stays_under_the_calibrated_line(p, group, k, gmp_fixtures):
    """findSignalSize (gpuLucas.cu:1064-1187): the measured choice lies in
    the candidate window, at or under the calibrated bits-per-word line, and
    its residue equals GMP's."""
    info = lucas.autotune_N(p)
    N = info["N"]
    print(f"p={p} chose N={N} ({p / N:.3f} bits/word)", flush=True)
    assert N in lucas.autotune_candidates(p)
    assert p / N <= lucas.max_bits_per_word(N)
    with lucas.Context(p, N) as ctx:
        ctx.step(k)
        assert ctx.res64() == fixture_res64(gmp_fixtures, group, p, k)
        assert ctx.max_roundoff() < ROUNDOFF_LIMIT


def test_reference_checkpoint_resume_vs_gmp(tmp_path, gmp_fixtures):
    """A run checkpointed in the reference's format (as gpuLucas writes it
    after loop index 1001, i.e. s_1000) resumes here: GMP's residue at
    10000, also when only the backup file is left."""
    from refckpt import write_reference_checkpoint
    p, N = 86243, 4608
    with lucas.Context(p, N) as ctx:
        ctx.step(1000)
        x = ctx.digits()
    path = str(tmp_path / "86243.chk")
    write_reference_checkpoint(path, p, N, 1001, x)
    want, zero = fixture_res64(gmp_fixtures, "c1_p86243", p, 10000)
    with lucas.Context.load(path) as ctx:
        assert ctx.iteration == 1000 and ctx.N == N
        ctx.step(9000)
        assert ctx.res64() == (want, zero)
    import os
    os.rename(path, path + ".bak")
    with lucas.Context.load(path) as ctx:
        assert ctx.iteration == 1000
        ctx.step(9000)
        assert ctx.res64() == (want, zero)


def test_checkpoint_resume_identical(tmp_path):
    p, N = 86243, 4608
    path = str(tmp_path / "m86243.chk")
    with lucas.Context(p, N) as ctx:
        ctx.step(3000)
        ctx.save(path)
        ctx.step(2000)
        want = ctx.res64()
    with lucas.Context.load(path) as ctx2:
        assert ctx2.iteration == 3000 and ctx2.N == N
        ctx2.step(2000)
        assert ctx2.res64() == want
    # full test with periodic checkpoints then resume from the file
    r1 = lucas.mersenne_test(p, N, checkpoint_every=20000, checkpoint_path=path)
    assert r1["is_prime"] == 1


def test_prepare_rehearses_and_restores_the_state(gmp_fixtures):
    """lucas_ctx_prepare builds and runs a step once, then puts the state
    back: digits, iteration count and accumulators are unchanged, and the
    next step reaches GMP's residue."""
    p, N = 86243, 4608
    with lucas.Context(p, N) as ctx:
        ctx.step(600)
        x = ctx.digits()
        err = ctx.max_roundoff()
        bits = ctx.product_bits()
        for k in (20, 400):
            ctx.prepare(k)
            assert ctx.iteration == 600
            assert np.array_equal(ctx.digits(), x)
            assert ctx.max_roundoff() == err and ctx.product_bits() == bits
        ctx.step(400)
        assert ctx.res64() == fixture_res64(gmp_fixtures, "c1_p86243", p, 1000)


@pytest.mark.parametrize("p,N,A,K,iters,kind", [
    (299843, 16384, 64, 2, 300, 1),      # E = 1 columns and rows
    (449741, 24576, 192, 4, 300, 1),     # radix-3 columns
    (1049413, 57344, 448, 2, 300, 1),    # radix-7 columns
    (2976221, 163840, 256, 2, 300, 1),   # the headline shape: radix-4 columns, radix-5 rows
    (2976221, 163840, 320, 4, 300, 2),   # radix-5 columns, radix-4 rows; non-temporal spectra
    (9068767, 524288, 512, 2, 200, 1),   # E = 8
    (37156667, 2097152, 1024, 4, 100, 2),  # E = 16, carry depth 3, N = 2^21, non-temporal spectra
    (37156667, 2097152, 1024, 2, 100, 1),
    (37156667, 2097152, 1024, 4, 100, 4),  # rows split over two waves, non-temporal spectra
    (37156667, 2097152, 512, 4, 100, 3),   # rows of 2048 split over two waves
    (18984731, 1048576, 512, 4, 100, 4),
    (9068767, 524288, 512, 2, 100, 3),     # rows of 512 split: two 256-word halves
    (9068767, 524288, 128, 2, 100, 3),     # rows of 2048 split, E = 2 columns
    (37156667, 2097152, 1024, 4, 100, 8),  # rows and columns split over two waves, non-temporal spectra
    (37156667, 2097152, 1024, 2, 100, 5),  # columns split
    (18984731, 1048576, 1024, 4, 100, 6),  # columns split, one wave per row of 512
])
def test_line_kernels_equal_tiled_kernels(p, N, A, K, iters, kind):
    """The line kernels and the tiled kernels reach the same digits after
    `iters` iterations from s0 = 4 (both restate the reference's
    restoration, so their digit vectors are the same integers)."""
    print(f"case p={p} N={N} line{kind} A={A} K={K} iters={iters}", flush=True)
    from gpulucas_amd import tune
    in_table = any(r[0] == N and r[7] == 0 for r in tune.tuned_entries())
    with lucas.Context(p, N, plan=tiled_plan(p, N) if in_table else lucas.plan_list(p, N)[0]) as ctx:
        assert ctx.info["line"] == 0
        ctx.step(iters)
        want = ctx.digits()
    with lucas.Context(p, N, plan={"p": p, "N": N, "A": A, "C": K, "line": kind}) as ctx:
        assert (ctx.info["line"], ctx.info["A"], ctx.info["C"]) == (kind, A, K)
        ctx.step(iters)
        assert np.array_equal(ctx.digits(), want)
        assert 0.0 < ctx.max_roundoff() < ROUNDOFF_LIMIT


@pytest.mark.gpu
def test_autotune_shape_picks_a_plan_with_the_same_digits():
    """lucas_autotune_shape at N = 2^19 (line shapes with and without
    non-temporal spectra first, then tilings): the plan it keeps reaches the
    default plan's digits, and its time is a positive ms/iteration."""
    p, N = 9068767, 524288
    print(f"case autotune_shape p={p} N={N}", flush=True)
    info, ms = lucas.autotune_shape(p, N, iters=40, max_candidates=8)
    assert info["N"] == N and 0.0 < ms < 1.0
    with lucas.Context(p, N) as ctx:
        ctx.step(50)
        want = ctx.digits()
    with lucas.Context(p, N, plan=info) as ctx:
        assert (ctx.info["A"], ctx.info["C"], ctx.info["line"]) == (info["A"], info["C"], info["line"])
        ctx.step(50)
        assert np.array_equal(ctx.digits(), want)


@pytest.mark.parametrize("p,N,A,K,kind", [(141731, 8192, 64, 2, 1), (2267549, 131072, 1024, 4, 1),
                                          (599657, 32768, 256, 1, 1), (2398633, 131072, 64, 2, 3),
                                          (4535099, 262144, 64, 2, 4), (2267549, 131072, 1024, 4, 5)])
def test_line_kernels_vs_oracle_steady_state(p, N, A, K, kind):
    """Line kernels from random digits: one and three iterations against the
    oracle's restatement (carry depth 3, E = 16 columns, one own column,
    rows of 1024 and 2048 and columns of 1024 split over two waves)."""
    print(f"case p={p} N={N} line{kind} A={A} K={K}", flush=True)
    x = random_balanced(p, N, 17)
    with lucas.Context(p, N, plan={"p": p, "N": N, "A": A, "C": K, "line": kind}) as ctx:
        assert ctx.info["line"] == kind
        ctx.set_digits(x, 0)
        ctx.step(1)
        y, _ = oracle.ibdwt_iterate(p, N, x, 1)
        assert np.array_equal(ctx.digits(), y)
        ctx.step(2)
        y, _ = oracle.ibdwt_iterate(p, N, y, 2)
        assert np.array_equal(ctx.digits(), y)


def test_headline_and_long_transform_plans_are_line_plans():
    """The planner's single-exponent plans at the headline and 2^21 shapes
    are the line kernels (the tuned table's entries)."""
    assert lucas.plan(2976221, 163840)["line"] >= 1
    assert lucas.plan(37156667, 2097152)["line"] >= 1
