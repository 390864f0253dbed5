"""ORACLE -- test infrastructure only.

ctypes wrappers over the CPU oracles built by ``make -C oracle``:

* libgmp_ll.so       -- GMP Lucas-Lehmer residues (oracle/gmp_ll.c)
* libibdwt_oracle.so -- plain-C restatement of the reference's IBDWT
                        iteration (oracle/ibdwt_oracle.c)

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use
this package, and only as the checker; the product library never calls it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_GMP = None
_IBD = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        build()
    return C.CDLL(path)


def _gmp():
    global _GMP
    if _GMP is None:
        lib = _load("libgmp_ll.so")
        lib.gmp_ll_res64.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        lib.gmp_ll_res64.restype = C.c_int
        lib.gmp_ll_residue_words.argtypes = [C.c_int, C.c_longlong, C.c_void_p, C.c_int]
        lib.gmp_ll_residue_words.restype = C.c_int
        lib.gmp_digits_res64.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        lib.gmp_digits_res64.restype = C.c_int
        lib.gmp_ll_time.argtypes = [C.c_int, C.c_int, C.c_int]
        lib.gmp_ll_time.restype = C.c_double
        lib.gmp_ll_version.restype = C.c_char_p
        _GMP = lib
    return _GMP


def _ibd():
    global _IBD
    if _IBD is None:
        lib = _load("libibdwt_oracle.so")
        lib.ibdwt_iterate.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_longlong, C.c_int, C.c_int,
                                      C.POINTER(C.c_double)]
        lib.ibdwt_conv_exact.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        lib.ibdwt_conv_fft.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]
        lib.ibdwt_restore.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ibdwt_res64.argtypes = [C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_int)]
        lib.ibdwt_widths_double.argtypes = [C.c_int, C.c_int, C.c_void_p]
        lib.ibdwt_widths_double.restype = None
        lib.ibdwt_tables.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ibdwt_carry_depth.argtypes = [C.c_int]
        lib.ibdwt_product_bits.argtypes = [C.c_int, C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.ibdwt_product_bits.restype = None
        _IBD = lib
    return _IBD


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


# ---- GMP ----------------------------------------------------------------------

def gmp_res64(p: int, iters) -> list[tuple[int, bool]]:
    its = np.ascontiguousarray(sorted(iters), dtype=np.int64)
    out = np.zeros(len(its), dtype=np.uint64)
    z = np.zeros(len(its), dtype=np.int32)
    rc = _gmp().gmp_ll_res64(p, _p(its), len(its), _p(out), _p(z))
    assert rc == 0
    return [(int(o), bool(zz)) for o, zz in zip(out, z)]


def gmp_residue(p: int, k: int) -> int:
    nw = (p + 31) // 32
    w = np.zeros(nw, dtype=np.uint32)
    rc = _gmp().gmp_ll_residue_words(p, k, _p(w), nw)
    assert rc >= 0
    return sum(int(x) << (32 * i) for i, x in enumerate(w))


def gmp_digits_res64(p: int, digits, offsets) -> tuple[int, bool]:
    d = np.ascontiguousarray(digits, dtype=np.int64)
    e = np.ascontiguousarray(offsets, dtype=np.int64)
    r = np.zeros(1, dtype=np.uint64)
    z = np.zeros(1, dtype=np.int32)
    assert _gmp().gmp_digits_res64(p, _p(d), _p(e), len(d), _p(r), _p(z)) == 0
    return int(r[0]), bool(z[0])


def gmp_time_ms(p: int, warm: int, iters: int) -> float:
    return _gmp().gmp_ll_time(p, warm, iters)


def gmp_version() -> str:
    return _gmp().gmp_ll_version().decode()


# ---- IBDWT restatement ------------------------------------------------------------

def s0_digits(N: int) -> np.ndarray:
    x = np.zeros(N, dtype=np.int32)
    x[0] = 4
    return x


def ibdwt_iterate(p: int, N: int, x, iters: int, exact: bool = False, mutation: int = 0):
    """Returns (digits after `iters` iterations, max roundoff)."""
    xx = np.array(x, dtype=np.int32, copy=True)
    err = C.c_double(0)
    rc = _ibd().ibdwt_iterate(p, N, _p(xx), iters, int(exact), mutation, C.byref(err))
    assert rc == 0, "oracle plan refused (q >= 21?)"
    return xx, err.value


def ibdwt_conv_exact(p: int, N: int, x) -> np.ndarray:
    xx = np.ascontiguousarray(x, dtype=np.int32)
    c = np.zeros(N, dtype=np.int64)
    assert _ibd().ibdwt_conv_exact(p, N, _p(xx), _p(c)) == 0
    return c


def ibdwt_conv_fft(p: int, N: int, x):
    xx = np.ascontiguousarray(x, dtype=np.int32)
    c = np.zeros(N, dtype=np.int64)
    err = C.c_double(0)
    assert _ibd().ibdwt_conv_fft(p, N, _p(xx), _p(c), C.byref(err)) == 0
    return c, err.value


def product_bits(c) -> tuple[int, int]:
    """(largest bit length of |c_j|, largest 64-bit leading-zero count of
    |c_j|) -- computeMaxBitVector + findMaxErrorHOST."""
    cc = np.ascontiguousarray(c, dtype=np.int64)
    b, z = C.c_int(0), C.c_int(0)
    _ibd().ibdwt_product_bits(len(cc), _p(cc), C.byref(b), C.byref(z))
    return b.value, z.value


def ibdwt_restore(p: int, N: int, c):
    cc = np.ascontiguousarray(c, dtype=np.int64)
    d = np.zeros(N, dtype=np.int32)
    h = np.zeros(N, dtype=np.int8)
    assert _ibd().ibdwt_restore(p, N, _p(cc), _p(d), _p(h)) == 0
    return d, h


def ibdwt_res64(p: int, N: int, x) -> tuple[int, bool]:
    xx = np.ascontiguousarray(x, dtype=np.int32)
    r = C.c_uint64(0)
    z = C.c_int(0)
    assert _ibd().ibdwt_res64(p, N, _p(xx), C.byref(r), C.byref(z)) == 0
    return r.value, bool(z.value)


def widths_double(p: int, N: int) -> np.ndarray:
    w = np.zeros(N, dtype=np.int32)
    _ibd().ibdwt_widths_double(p, N, _p(w))
    return w


def tables(p: int, N: int):
    w = np.zeros(N, dtype=np.int32)
    m = np.zeros(N, dtype=np.uint8)
    a = np.zeros(N, dtype=np.float64)
    ai = np.zeros(N, dtype=np.float64)
    assert _ibd().ibdwt_tables(p, N, _p(w), _p(m), _p(a), _p(ai)) == 0
    return w, m, a, ai


def carry_depth(q: int) -> int:
    return _ibd().ibdwt_carry_depth(q)


def offsets(p: int, N: int) -> np.ndarray:
    """e_j = ceil(j p / N) for j < N."""
    j = np.arange(N, dtype=np.int64)
    return (j * p + N - 1) // N
