"""Host-side logic of libgpulucas (no GPU needed): exports, planning, the
per-digit tables the kernels regenerate, and residue normalisation."""
import os
import re

import numpy as np
import pytest

import oracle
from gpulucas_amd import lucas, workloads as wl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "gpulucas.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lucas_[A-Za-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_function():
    import ctypes
    lib = ctypes.CDLL(lucas.LIB_PATH)
    names = header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    # the ctypes mirror binds exactly the declared set
    assert lucas.exported_symbols() == names
    # the profiling build exports the same set
    from gpulucas_amd import build
    plib = ctypes.CDLL(build.PHASES_LIB)
    for n in names:
        assert hasattr(plib, n), n


def test_phase_stamps_need_the_profiling_build():
    # the default library refuses before touching a device; bad handles first
    f = lucas._lib().lucas_ctx_phase_stamps
    assert f(None, 1, None, None, 0, None, None) == lucas.ERR_INVALID_ARG
    assert lucas._lib().lucas_status_string(lucas.ERR_UNSUPPORTED) == b"not supported by this build"


def test_phase_summary_of_stamps():
    import numpy as np
    from gpulucas_amd import phases
    # two workgroups; ticks of 10 ns
    st = np.array([[100, 300, 500, 600, 900, 910],
                   [150, 350, 600, 700, 1000, 1030]], np.int64)
    s = phases.summarize(st, phases.COL_PHASES)
    assert s["workgroups"] == 2
    assert s["span_us"] == 9.3 and s["start_spread_us"] == 0.5
    assert s["workgroup_us_mean"] == 8.45 and s["workgroup_us_max"] == 8.8
    assert s["phase_us_mean"]["load"] == 2.0
    assert s["phase_us_mean"]["inverse FFT + unweight"] == 2.25
    assert s["phase_us_max"]["store"] == 0.3
    bad = st.copy()
    bad[0, 3] = 400  # carry ends before it starts
    with pytest.raises(ValueError):
        phases.summarize(bad, phases.COL_PHASES)


def test_version_and_status_strings():
    assert "gfx950" in lucas.version()
    assert lucas._lib().lucas_status_string(lucas.ERR_ROUNDOFF) == b"roundoff limit exceeded"


@pytest.mark.parametrize("p,N,q,n", [(86243, 4608, 18, 2), (1257787, 65536, 19, 2), (2976221, 163840, 18, 2),
                                     (859433, 49152, 17, 3), (1279, 64, 19, 2), (4423, 240, 18, 2)])
def test_plan_config_shapes(p, N, q, n):
    info = lucas.plan(p, N)
    assert info["N"] == N and info["q"] == q and info["carry_depth"] == n
    assert info["A"] * info["B"] == N // 2
    assert info["n_wide"] == p % N
    assert info["halo"] == (n + 2) // 2
    assert info["B"] % info["C"] == 0
    assert info["col_wgs"] == info["B"] // info["C"]
    assert info["row_wgs"] == (info["A"] + 1) // 2
    assert info["device_bytes"] >= 20 * N


def test_plan_rejects_bad_inputs():
    for p, N in [(2976221, 163841), (2976221, 11 * 16384), (2976221, 131072), (5, 4), (86243, 0 - 2)]:
        with pytest.raises(lucas.LucasError):
            lucas.plan(p, N)


@pytest.mark.parametrize("A", [40, 64, 80, 128, 160, 256, 320, 512, 640, 1024])
def test_plan_forced_shapes(A):
    info = lucas.plan(2976221, 163840, A=A)
    assert info["A"] == A and A * info["B"] == 81920


@pytest.mark.parametrize("p", [86243, 1257787, 2976221, 4990753, 1062497, 39999983, 110503])
def test_default_length(p):
    N = lucas.default_N(p)
    info = lucas.plan(p, N)
    assert p / N <= 20.2
    assert info["q"] <= 20
    # smallest admissible: the next smaller smooth length exceeds the rule
    assert N >= p / 20.2


@pytest.mark.parametrize("p,N", [(100000007, 6291456), (136279841, 8064000)])
def test_wide_tilings_extend_the_longest_lengths(p, N):
    """Past the ordinary tilings (at most 4 words per work-item) the planner
    offers wide ones, whose Stockham stages of radix 3, 5, 7 and 8 hold up to
    8 words per work-item: exponents up to M52 = 2^136279841 - 1."""
    assert lucas.default_N(p) == N
    plans = lucas.plan_list(p, N)
    assert plans
    for info in plans:
        T = info["threads"]
        tile = max(info["A"] * (info["C"] + info["halo"]), 2 * info["B"])
        assert 4 * T < tile <= 8 * T
        # a wave factor holds at most 4 words per work-item: wide column tiles have none
        if info["A"] * (info["C"] + info["halo"]) > 4 * T:
            assert info["wave_A"] == 0
        if 2 * info["B"] > 4 * T:
            assert info["wave_B"] == 0
    assert lucas.plan(p, N) in plans


def test_reference_checkpoint_format_is_read(tmp_path):
    """A checkpoint written by the reference (weighted double signal) is
    parsed back to its digits; damaged ones are refused before any device
    work."""
    from refckpt import write_reference_checkpoint
    p, N = 86243, 4608
    x, _ = oracle.ibdwt_iterate(p, N, oracle.s0_digits(N), 5)
    good = str(tmp_path / "86243.chk")
    write_reference_checkpoint(good, p, N, 6, x)  # loop index 6 holds s_5
    try:
        ctx = lucas.Context.load(good)
    except lucas.LucasError as e:
        assert e.status == lucas.ERR_NO_DEVICE  # parsed; no GPU in this process
    else:
        with ctx:
            assert ctx.iteration == 5 and np.array_equal(ctx.digits(), x)
    bad = str(tmp_path / "bad.chk")
    # truncated signal
    with open(good, "rb") as f:
        blob = f.read()
    with open(bad, "wb") as f:
        f.write(blob[:-8])
    with pytest.raises(lucas.LucasError) as e:
        lucas.Context.load(bad)
    assert e.value.status == lucas.ERR_IO
    # a value between two digits is not a state
    y = x.astype(np.float64)
    y[7] += 0.5
    write_reference_checkpoint(bad, p, N, 6, y)
    with pytest.raises(lucas.LucasError) as e:
        lucas.Context.load(bad)
    assert e.value.status == lucas.ERR_IO
    # another program's file
    write_reference_checkpoint(bad, p, N, 6, x, name=b"CUDALucas")
    with pytest.raises(lucas.LucasError) as e:
        lucas.Context.load(bad)
    assert e.value.status == lucas.ERR_IO


def test_smallest_supported_exponent():
    # below p = 32 no length leaves room for the carry restoration and the halo
    assert lucas.default_N(31) == 0
    with pytest.raises(lucas.LucasError):
        lucas.plan(31, 0)
    assert lucas.default_N(37) == 16


def test_longest_supported_exponent():
    # beyond N = 8064000 no two-pass tiling fits: refused, not mis-planned
    assert lucas.default_N(150000001) == 0
    with pytest.raises(lucas.LucasError):
        lucas.plan(150000001, 0)
    # the ordinary range keeps its lengths
    assert lucas.default_N(82589933) == 4915200


@pytest.mark.parametrize("p,N", [(86243, 4608), (2976221, 163840), (1257787, 65536), (4423, 240)])
def test_digit_tables_match_oracle(p, N):
    w, a, ai = lucas.digit_tables(p, N)
    ow, _, oa, oai = oracle.tables(p, N)
    assert np.array_equal(w, ow)
    assert np.array_equal(a, oa)
    assert np.max(np.abs(ai - oai) / oai) < 2.3e-16


@pytest.mark.parametrize("p,N", [(1279, 64), (4423, 240), (86243, 4608)])
def test_digits_res64_matches_gmp(p, N):
    rng = np.random.default_rng(p)
    e = oracle.offsets(p, N)
    w, _, _ = lucas.digit_tables(p, N)
    for trial in range(5):
        lim = (1 << (w + 1)).astype(np.int64)
        d = (rng.integers(0, 2 * lim) - lim).astype(np.int32)
        r, z, norm = lucas.digits_res64(p, N, d)
        assert (r, z) == oracle.gmp_digits_res64(p, d, e)
        assert np.all(norm >= 0) and np.all(norm < (1 << w))
        assert oracle.gmp_digits_res64(p, norm, e) == (r, z)


@pytest.mark.parametrize("p,N", [(1279, 64), (4423, 240)])
def test_digits_res64_zero_forms(p, N):
    w, _, _ = lucas.digit_tables(p, N)
    zero = np.zeros(N, dtype=np.int32)
    assert lucas.digits_res64(p, N, zero)[:2] == (0, True)
    ones = ((1 << w) - 1).astype(np.int32)  # = M_p
    assert lucas.digits_res64(p, N, ones)[:2] == (0, True)
    neg = zero.copy()
    neg[0] = -1  # -1 == M_p - 1: res64 all ones in the low word
    r, z, _ = lucas.digits_res64(p, N, neg)
    assert not z and r == (2 ** p - 2) & (2 ** 64 - 1)


def test_workload_lists_are_deterministic_primes():
    a = wl.batch_primes()
    assert a == wl.batch_primes() and len(set(a)) == 64
    assert all(wl.is_prime(p) and 1_000_000 <= p <= 5_000_000 for p in a)
    b = wl.near3m_primes()
    assert all(2_900_000 <= p <= 3_000_000 and wl.is_prime(p) for p in b)
    jobs = wl.serving_jobs(count=32)
    assert all(t2 >= t1 for (t1, _, _), (t2, _, _) in zip(jobs, jobs[1:]))
    assert all(1e5 <= p <= 4.1e7 and 5000 <= k <= 50000 for _, p, k in jobs)


def test_miller_rabin_against_sieve():
    n = 20000
    sieve = np.ones(n, dtype=bool)
    sieve[:2] = False
    for i in range(2, int(n ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i::i] = False
    assert [i for i in range(n) if wl.is_prime(i)] == list(np.nonzero(sieve)[0])


def test_tuned_table_entries_are_the_plans_chosen():
    """Every measured shape is admissible and is what the planner returns for
    the exponent it was measured at: a length's line entry for a single
    exponent, its tiling otherwise (and always for a batch's split)."""
    from gpulucas_amd import tune
    table = tune.tuned_entries()
    assert len(table) >= 40
    assert len({(row[0], row[7]) for row in table}) == len(table)
    line_n = {row[0] for row in table if row[7]}
    assert {row[0] for row in table if not row[7]} >= line_n, "every line entry keeps its tiling for batches"
    for N, A, C, T, pref, wA, wB, line, p in table:
        if line:
            pl = lucas.plan(p, N)
            assert (pl["line"], pl["A"], pl["C"], pl["threads"]) == (line, A, C, T), N
        elif N not in line_n:
            pl = lucas.plan(p, N)
            assert pl["line"] == 0
            assert (pl["A"], pl["C"], pl["threads"], pl["radix_pref"]) == (A, C, T, pref), N
            assert (pl["wave_A"], pl["wave_B"]) == (wA, wB), N


@pytest.mark.parametrize("p", [86243, 1257787, 2976221, 37156667])
def test_autotune_candidates_follow_the_reference_window(p):
    """findSignalSize's window (gpuLucas.cu:1072-1077): N from p/20 to the
    smallest power of two >= p/16, but not past p/14.  Every candidate is an
    admissible length at or under the calibrated bits-per-word line, every
    admissible length of the window under the line is a candidate, and the
    default length is always among them."""
    from gpulucas_amd import lucas
    lo = -(-p // 20)
    pow2 = 1
    while pow2 < p // 16:
        pow2 *= 2
    hi = min(pow2, p // 14)
    cands = lucas.autotune_candidates(p)
    assert cands == sorted(set(cands))
    dn = lucas.default_N(p)
    assert dn in cands
    for N in cands:
        assert N == dn or lo <= N <= hi
        assert p / N <= lucas.max_bits_per_word(N)
        lucas.plan(p, N)  # has a tiling
    smooth = [N for N in range(lo + (lo & 1), hi + 1, 2) if _smooth(N) and p / N <= lucas.max_bits_per_word(N)]
    for N in smooth:
        try:
            lucas.plan(p, N)
        except lucas.LucasError:
            continue
        assert N in cands, N


def _smooth(n):
    for f in (2, 3, 5, 7):
        while n % f == 0:
            n //= f
    return n == 1
