"""CPU emulation of the engine's kernels (tests only)."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(HERE, "libemu.so")
SOURCES = [os.path.join(ROOT, "gpulucas_amd", "csrc", "lucas_kernels.hip"),
           os.path.join(ROOT, "gpulucas_amd", "csrc", "lucas_line.hip"),
           os.path.join(HERE, "emu_driver.cpp"),
           os.path.join(ROOT, "gpulucas_amd", "csrc", "lucas_host.cpp")]
_lib = None


def build():
    # each process builds into its own file and renames it into place, so
    # parallel test workers never load a half-written library
    tmp = f"{LIB}.{os.getpid()}.tmp"
    cmd = ["g++", "-std=c++20", "-O2", "-fPIC", "-shared", "-pthread", "-I" + HERE,
           "-I" + os.path.join(ROOT, "gpulucas_amd", "csrc"), "-I" + os.path.join(ROOT, "include"),
           "-x", "c++", SOURCES[0], "-x", "c++", SOURCES[1], "-x", "c++", SOURCES[2], SOURCES[3], "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)


def lib():
    global _lib
    if _lib is None:
        headers = [os.path.join(HERE, "hip", "hip_runtime.h")] + [
            os.path.join(ROOT, "gpulucas_amd", "csrc", h)
            for h in ("lucas_desc.h", "lucas_host.h", "lucas_device.h", "lucas_line.h")]
        newest = max(os.path.getmtime(s) for s in SOURCES + headers)
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < newest:
            build()
        _lib = C.CDLL(LIB)
        _lib.emu_set_single.argtypes = [C.c_int]
        _lib.emu_set_waves.argtypes = [C.c_int, C.c_int]
        _lib.emu_run.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_longlong, C.c_void_p,
                                 C.POINTER(C.c_double)]
        _lib.emu_product_bits.argtypes = [C.c_void_p]
        _lib.emu_set_line.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    return _lib


def run(p, N, digits, iters, A=0, C_=0, T=0, single=False, waves=(-1, -1), line=None):
    """Digits after `iters` iterations of the emulated kernels; `single` uses
    the single-job kernels (descriptor by value) instead of the batched ones;
    `waves` forces the column / row wave factors (-1: planner's choice);
    `line` = (A, K[, split[, col_split]]) runs the line kernels with A rows,
    K own columns per column workgroup (0: first instantiated choice) and
    rows / columns held by one wave (0) or two (1; -1 or absent: the first
    choice)."""
    lib().emu_set_single(1 if single else 0)
    lib().emu_set_line(*((tuple(line) + (-1, -1))[:4] if line is not None else (-1, 0, -1, -1)))
    lib().emu_set_waves(*waves)
    x = np.ascontiguousarray(digits, dtype=np.int32)
    out = np.zeros(N, dtype=np.int32)
    err = C.c_double(0)
    rc = lib().emu_run(p, N, A, C_, T, x.ctypes.data_as(C.c_void_p), iters, out.ctypes.data_as(C.c_void_p),
                       C.byref(err))
    assert rc == 0, "plan refused"
    return out, err.value


def product_bits():
    """(max bit length, max 64-bit leading-zero count) of the rounded
    products over the last run's iterations (the kernels' diagnostic)."""
    out = np.zeros(2, dtype=np.int32)
    lib().emu_product_bits(out.ctypes.data_as(C.c_void_p))
    return int(out[0]), int(out[1])
