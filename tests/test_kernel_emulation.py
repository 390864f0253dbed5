"""The engine's kernel source, compiled for the host under a SIMT emulation
(tests/emu), must reproduce the oracle's digits bit for bit.  This checks the
kernels' indexing, LDS staging, halo wrap and carry on the CPU."""
import numpy as np
import pytest

import emu
import oracle


@pytest.mark.parametrize("p,N,A,C,iters,T", [
    (1279, 64, 4, 1, 3, 0),       # C < halo: halo spans a previous tile and a row wrap
    (1279, 64, 8, 4, 3, 0),       # single column tile
    (4423, 240, 15, 8, 2, 0),     # odd A (radix 3, 5), one tile
    (9941, 560, 7, 4, 2, 0),      # radix 7 columns
    (2203, 120, 0, 0, 2, 0),      # planner's own choice
    (44497, 2560, 40, 16, 2, 256),  # A*C > T: the carry runs in three passes
    (294919, 16384, 64, 4, 1, 512),    # wave factors only: columns 64 (P=1), rows 128 (P=2)
    (218737, 12288, 192, 2, 1, 1024),  # columns 3 x 64: Stockham radix 3 then wave factor
    (294919, 16384, 512, 2, 1, 512),   # columns 2 x 256 (P=4)
    (368647, 20480, 16, 8, 1, 1024),   # rows 5 x 128: radix 5 then wave factor
])
def test_emulated_kernels_match_oracle(p, N, A, C, iters, T):
    x = oracle.s0_digits(N)
    got, err = emu.run(p, N, x, iters, A, C, T)
    want, _ = oracle.ibdwt_iterate(p, N, x, iters)
    assert np.array_equal(got, want)
    assert err < 0.1


def test_emulated_kernels_steady_state():
    p, N = 4423, 240
    rng = np.random.default_rng(5)
    w, _, _, _ = oracle.tables(p, N)
    half = (1 << (w - 1)).astype(np.int64)
    x = (rng.integers(0, 2 * half) - half).astype(np.int32)
    got, _ = emu.run(p, N, x, 2, 15, 8)
    want, _ = oracle.ibdwt_iterate(p, N, x, 2)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("p,N,A,C,T,wA,wB", [
    (294919, 16384, 64, 4, 512, 64, 128),
    (218737, 12288, 192, 2, 1024, 64, 0),
    (294919, 16384, 512, 2, 512, 256, 0),
    (368647, 20480, 16, 8, 1024, 0, 128),
])
def test_emulated_shapes_take_the_wave_path(p, N, A, C, T, wA, wB):
    """The wave-factor cases above really run the wave-register transforms."""
    from gpulucas_amd import lucas
    pl = [x for x in lucas.plan_list(p, N) if x["A"] == A and x["C"] == C and x["threads"] == T][0]
    assert (pl["wave_A"], pl["wave_B"]) == (wA, wB)


@pytest.mark.parametrize("p,N,A,C,iters,T", [
    (1279, 64, 4, 1, 2, 0),          # C < halo, row wrap
    (9941, 560, 7, 4, 2, 0),         # odd A, radix 7
    (44497, 2560, 40, 1, 1, 256),    # 32 column tiles: XCD-dealt tile order
    (294919, 16384, 64, 4, 1, 512),  # wave factors, 32 column tiles
])
def test_emulated_single_job_kernels_match_oracle(p, N, A, C, iters, T):
    """The single-job kernels (descriptor as kernel argument, tile from the
    workgroup id and XCD dealing) compute the same digits."""
    x = oracle.s0_digits(N) if iters > 1 else None
    if x is None:
        w, _, _, _ = oracle.tables(p, N)
        rng = np.random.default_rng(7)
        half = (1 << (w - 1)).astype(np.int64)
        x = (rng.integers(0, 2 * half) - half).astype(np.int32)
    got, err = emu.run(p, N, x, iters, A, C, T, single=True)
    want, _ = oracle.ibdwt_iterate(p, N, x, iters)
    assert np.array_equal(got, want)
    assert err < 0.4


@pytest.mark.parametrize("p,N,A,C,T,waves", [
    (164729, 8192, 256, 2, 1024, (64, 0)),    # radix-4 Stockham stage, then 64-point wave factor
    (164729, 8192, 256, 2, 1024, (128, 0)),   # radix 2, then 128
    (164729, 8192, 256, 4, 1024, (0, 0)),     # Stockham only
    (368647, 20480, 16, 4, 1024, (0, 0)),     # rows Stockham only (5 x 128 as radices)
])
def test_emulated_smaller_wave_factors_match_oracle(p, N, A, C, T, waves):
    """Every wave factor the planner offers (not only the largest) computes
    the oracle's digits."""
    w, _, _, _ = oracle.tables(p, N)
    rng = np.random.default_rng(9)
    half = (1 << (w - 1)).astype(np.int64)
    x = (rng.integers(0, 2 * half) - half).astype(np.int32)
    got, err = emu.run(p, N, x, 1, A, C, T, single=True, waves=waves)
    want, _ = oracle.ibdwt_iterate(p, N, x, 1)
    assert np.array_equal(got, want)
    assert err < 0.4


@pytest.mark.parametrize("p,N,A,C", [
    (294919, 16384, 512, 1),  # column tile 1536 words on 256 items: radix 8.8.8 stages, 6 words each
    (180181, 10240, 320, 2),  # column tile 1280: radix 8.8.5, 5 words per item
    (322561, 18432, 16, 32),  # row pair 1152 words: radix 8.8.3.3, 6 words per item
])
def test_emulated_wide_tiles_match_oracle(p, N, A, C):
    """Wide tilings (more than 4 words per work-item, which only radix 3, 5,
    7 and 8 stages hold) compute the oracle's digits."""
    from gpulucas_amd import lucas
    info = lucas.plan(p, N, A=A, C_=C)  # not an ordinary tiling at 256 items
    assert max(A * (C + info["halo"]), 2 * info["B"]) > 4 * 256
    x = oracle.s0_digits(N)
    got, err = emu.run(p, N, x, 2, A, C, 256, single=True, waves=(0, 0))
    want, _ = oracle.ibdwt_iterate(p, N, x, 2)
    assert np.array_equal(got, want)
    assert err < 0.4


@pytest.mark.parametrize("p,N,A,C,iters", [
    (4423, 240, 15, 8, 1),   # odd A, one tile
    (9941, 560, 7, 4, 3),    # radix 7, several iterations: the maximum over all of them
    (1279, 64, 4, 1, 2),     # C < halo: halo products counted as well
])
def test_emulated_product_bits_match_exact_convolution(p, N, A, C, iters):
    """The kernels' product-size diagnostic (computeMaxBitVector) equals
    the bit lengths of the exact weighted convolution's outputs, maximised
    over every iteration of the run."""
    w, _, _, _ = oracle.tables(p, N)
    rng = np.random.default_rng(21)
    half = (1 << (w - 1)).astype(np.int64)
    x = (rng.integers(0, 2 * half) - half).astype(np.int32)
    want_b, want_z = 0, 0
    y = x
    for _ in range(iters):
        b, z = oracle.product_bits(oracle.ibdwt_conv_exact(p, N, y))
        want_b, want_z = max(want_b, b), max(want_z, z)
        y, _ = oracle.ibdwt_iterate(p, N, y, 1)
    got, _ = emu.run(p, N, x, iters, A, C, 0)
    assert np.array_equal(got, y)
    assert emu.product_bits() == (want_b, want_z)


# Line kernels: every row and column length instantiated (64E, E = 1, 2, 3,
# 4, 5, 7, 8, 16), carry depths 2 and 3, K = 1, 2, 4 own columns, the halo
# wrap of tile 0, one and several iterations, batched and single-job forms,
# rows of 1024 and 2048 words and columns of 1024 words split over two waves,
# and the N = 2^21 plan itself.
@pytest.mark.parametrize("p,N,A,K,iters,single,split", [
    (149921, 8192, 64, 2, 3, False, 0),     # A = B = 64 (E = 1): several iterations from random digits
    (141731, 8192, 64, 2, 1, True, 0),      # carry depth 3 (q = 17)
    (299843, 16384, 128, 2, 1, False, 0),   # columns E = 2, rows E = 2
    (449741, 24576, 192, 2, 1, True, 0),    # columns E = 3 (radix 3), rows E = 2
    (449741, 24576, 64, 4, 1, False, 0),    # rows E = 3
    (599657, 32768, 256, 1, 1, True, 0),    # columns E = 4, one own column per workgroup
    (749587, 40960, 320, 2, 2, False, 0),   # columns E = 5 (radix 5)
    (1049413, 57344, 448, 4, 1, True, 0),   # columns E = 7 (radix 7)
    (1199309, 65536, 512, 4, 1, False, 0),  # columns E = 8 (lanes hold one radix-8 butterfly each)
    (2267549, 131072, 1024, 4, 1, True, 0),   # columns E = 16 (two butterflies per lane), depth 3
    (2998279, 163840, 256, 2, 1, True, 0),    # rows E = 5 (the headline shape)
    (2098801, 114688, 128, 4, 1, False, 0),   # rows E = 7
    (2398633, 131072, 128, 4, 1, True, 0),    # rows E = 8
    (4797241, 262144, 128, 4, 1, True, 0),    # rows E = 16
    (2398633, 131072, 64, 2, 1, True, 1),   # rows of 1024 split: two 512-word halves (register twiddles)
    (4535099, 262144, 64, 2, 1, False, 1),  # rows of 2048 split: two 1024-word halves, carry depth 3
    (2267549, 131072, 1024, 4, 1, False, 2),  # columns of 1024 split: two 512-word halves, depth 3
    (2398633, 131072, 1024, 2, 2, True, 2),   # ... K = 2, depth 2, two iterations
    (36280733, 2097152, 1024, 4, 1, True, 3),  # N = 2^21 as planned: rows and columns split, depth 3
])
def test_emulated_line_kernels_match_oracle(p, N, A, K, iters, single, split):
    w, _, _, _ = oracle.tables(p, N)
    rng = np.random.default_rng(p % 1000)
    half = (1 << (w - 1)).astype(np.int64)
    x = (rng.integers(0, 2 * half) - half).astype(np.int32)
    # split: bit 0 rows split over two waves, bit 1 columns split
    line = (A, K, split & 1, split >> 1)
    got, err = emu.run(p, N, x, iters, single=single, line=line)
    want, _ = oracle.ibdwt_iterate(p, N, x, iters)
    assert np.array_equal(got, want)
    assert 0.0 < err < 0.1
