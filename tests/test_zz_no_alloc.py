"""The entry point that allocates no device memory
(lucas_step_digits_noalloc) against the context path on the same inputs,
through the C ABI: two calls in a row on one GPU, the second with a smaller
input, each bit-identical to lucas_ctx_step's digits and roundoff."""
import numpy as np
import pytest

import oracle
from gpulucas_amd import lucas

pytestmark = pytest.mark.gpu


def random_balanced(p, N, seed):
    rng = np.random.default_rng(seed)
    w, _, _ = lucas.digit_tables(p, N)
    half = (1 << (w - 1)).astype(np.int64)
    return (rng.integers(0, 2 * half) - half).astype(np.int32)


def _reference(p, N, x, iters):
    with lucas.Context(p, N) as ctx:
        ctx.set_digits(x, 0)
        ctx.step(iters)
        return ctx.digits(), ctx.max_roundoff()


def _noalloc(p, N, x, iters):
    # caller-owned device memory: a context's digit vector, written and read
    # through the context, advanced by the no-allocation entry point
    with lucas.Context(p, N) as holder:
        holder.set_digits(x, 0)
        err = lucas.step_digits_noalloc(p, N, holder.device_digits(), iters)
        return holder.digits(), err


def test_noalloc_matches_context_path_twice_second_smaller():
    cases = [(2976221, 163840, 40), (86243, 4608, 25)]
    for p, N, iters in cases:
        x = random_balanced(p, N, 41)
        want, want_err = _reference(p, N, x, iters)
        got, err = _noalloc(p, N, x, iters)
        print(f"p={p} N={N} iters={iters} roundoff {err:.4f}", flush=True)
        assert np.array_equal(got, want), (p, N)
        assert err == want_err and 0.0 < err < lucas.DEFAULT_ERROR_LIMIT


def test_noalloc_split_row_plan_matches_context_path():
    """N = 393216 plans the line kernels with rows split over two waves (its
    tuned kind 3): the scratch-array path reaches the context path's digits."""
    p, N, iters = 7149713, 393216, 20
    assert lucas.plan(p, N)["line"] == 3
    x = random_balanced(p, N, 43)
    want, want_err = _reference(p, N, x, iters)
    got, err = _noalloc(p, N, x, iters)
    print(f"p={p} N={N} iters={iters} roundoff {err:.4f}", flush=True)
    assert np.array_equal(got, want)
    assert err == want_err


def test_noalloc_from_s0_matches_oracle():
    p, N = 9941, 560
    x = oracle.s0_digits(N)
    got, _ = _noalloc(p, N, x, 30)
    want, _ = oracle.ibdwt_iterate(p, N, x, 30)
    assert np.array_equal(got, want)


def test_noalloc_rejects_lengths_past_its_scratch():
    N = 2 * lucas.NOALLOC_MAX_N
    with lucas.Context(37156667, 2097152) as holder:
        with pytest.raises(lucas.LucasError) as e:
            lucas.step_digits_noalloc(37156667, 2097152, holder.device_digits(), 1)
        assert e.value.status == lucas.ERR_INVALID_ARG
    assert N > lucas.NOALLOC_MAX_N
