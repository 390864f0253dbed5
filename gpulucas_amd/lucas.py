"""Python mirror of the engine's C ABI (include/gpulucas.h), via ctypes.

The reference's user-facing operations keep their names and meaning:

* ``mersenneTest(p, N=0)`` -- full Lucas-Lehmer test of 2^p - 1
  (gpuLucas.cu:1358-1551): p-2 squarings from s_0 = 4, returns the result
  record (prime flag, Res64, N, max roundoff, ms/iteration).
* ``errorTest(p, N=0, iterations=0)`` -- timing/roundoff trial
  (errorTrial, gpuLucas.cu:1192-1282): mean ms per iteration, or -1.0 when
  the roundoff reached the limit (the reference's return value).
* ``findSignalSize(p)`` -- measured choice of N (gpuLucas.cu:1064-1187).

Everything runs in libgpulucas.so (HIP kernels for gfx950); there is no CPU
fallback, and importing the wrapper without the built library raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GPULUCAS_LIB", os.path.join(_HERE, "libgpulucas.so"))

OK = 0
ERR_INVALID_ARG = -1
ERR_OOM = -2
ERR_ROUNDOFF = -3
ERR_DEVICE = -4
ERR_IO = -5
ERR_NO_DEVICE = -6
ERR_BUSY = -7
ERR_UNSUPPORTED = -8
DEFAULT_ERROR_LIMIT = 0.30


class LucasError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = _lib().lucas_status_string(status).decode() if _LIB is not None else str(status)
        super().__init__(f"{what}: {msg} ({status})" if what else f"{msg} ({status})")


class PlanInfo(C.Structure):
    _fields_ = [("p", C.c_int), ("N", C.c_int), ("q", C.c_int), ("n_wide", C.c_int),
                ("carry_depth", C.c_int), ("A", C.c_int), ("B", C.c_int), ("C", C.c_int),
                ("halo", C.c_int), ("col_wgs", C.c_int), ("row_wgs", C.c_int),
                ("threads", C.c_int), ("radix_pref", C.c_int), ("wave_A", C.c_int), ("wave_B", C.c_int), ("col_dense", C.c_int),
                ("line", C.c_int), ("device_bytes", C.c_size_t)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Result(C.Structure):
    _fields_ = [("p", C.c_int), ("N", C.c_int), ("is_prime", C.c_int), ("res64", C.c_uint64),
                ("iterations", C.c_longlong), ("max_roundoff", C.c_double), ("ms_per_iter", C.c_double),
                ("max_product_bits", C.c_int), ("max_product_clz", C.c_int)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["res64_hex"] = f"0x{self.res64:016x}"
        return d


class Options(C.Structure):
    _fields_ = [("device", C.c_int), ("error_limit", C.c_double), ("checkpoint_every", C.c_longlong),
                ("checkpoint_path", C.c_char_p), ("resume", C.c_int)]


class ServerOptions(C.Structure):
    _fields_ = [("device", C.c_int), ("hbm_budget_bytes", C.c_size_t), ("max_jobs", C.c_int),
                ("chunk_iters", C.c_int), ("checkpoint_dir", C.c_char_p), ("checkpoint_every", C.c_longlong),
                ("error_limit", C.c_double)]


class JobResult(C.Structure):
    _fields_ = [("id", C.c_longlong), ("p", C.c_int), ("N", C.c_int), ("iterations", C.c_longlong),
                ("done", C.c_int), ("status", C.c_int), ("res64", C.c_uint64), ("is_zero", C.c_int),
                ("max_roundoff", C.c_double), ("submit_s", C.c_double), ("start_s", C.c_double),
                ("finish_s", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_LIB = None

# name -> (restype, argtypes); every function declared in include/gpulucas.h
_P = C.POINTER
_SIGS = {
    "lucas_version": (C.c_char_p, []),
    "lucas_status_string": (C.c_char_p, [C.c_int]),
    "lucas_device_count": (C.c_int, [_P(C.c_int)]),
    "lucas_default_N": (C.c_int, [C.c_int]),
    "lucas_max_bits_per_word": (C.c_double, [C.c_int]),
    "lucas_device_copy_bench": (C.c_int, [C.c_size_t, C.c_int, C.c_int, _P(C.c_double)]),
    "lucas_plan": (C.c_int, [C.c_int, C.c_int, _P(PlanInfo)]),
    "lucas_plan_shape": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _P(PlanInfo)]),
    "lucas_plan_list": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_int, _P(C.c_int)]),
    "lucas_line_plan_list": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_int, _P(C.c_int)]),
    "lucas_autotune_shape": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _P(Options), _P(PlanInfo),
                                       _P(C.c_double)]),
    "lucas_digit_tables": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "lucas_digits_res64": (C.c_int, [C.c_int, C.c_int, C.c_void_p, _P(C.c_uint64), _P(C.c_int), C.c_void_p]),
    "lucas_ctx_create": (C.c_int, [C.c_int, C.c_int, C.c_int, _P(C.c_void_p)]),
    "lucas_ctx_create_shape": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P(C.c_void_p)]),
    "lucas_ctx_create_plan": (C.c_int, [_P(PlanInfo), C.c_int, _P(C.c_void_p)]),
    "lucas_ctx_destroy": (C.c_int, [C.c_void_p]),
    "lucas_ctx_info": (C.c_int, [C.c_void_p, _P(PlanInfo)]),
    "lucas_ctx_step": (C.c_int, [C.c_void_p, C.c_longlong]),
    "lucas_ctx_iteration": (C.c_longlong, [C.c_void_p]),
    "lucas_ctx_get_digits": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lucas_ctx_set_digits": (C.c_int, [C.c_void_p, C.c_void_p, C.c_longlong]),
    "lucas_ctx_device_digits": (C.c_int, [C.c_void_p, _P(C.c_void_p)]),
    "lucas_ctx_res64": (C.c_int, [C.c_void_p, _P(C.c_uint64), _P(C.c_int)]),
    "lucas_ctx_max_roundoff": (C.c_int, [C.c_void_p, C.c_int, _P(C.c_double)]),
    "lucas_ctx_product_bits": (C.c_int, [C.c_void_p, C.c_int, _P(C.c_int), _P(C.c_int)]),
    "lucas_ctx_prepare": (C.c_int, [C.c_void_p, C.c_longlong]),
    "lucas_ctx_time_kernels": (C.c_int, [C.c_void_p, C.c_int, _P(C.c_double), _P(C.c_double)]),
    "lucas_ctx_phase_stamps": (C.c_int, [C.c_void_p, C.c_int, _P(C.c_ulonglong), _P(C.c_ulonglong), C.c_int,
                                         _P(C.c_int), _P(C.c_int)]),
    "lucas_ctx_time_iterations": (C.c_int, [C.c_void_p, C.c_longlong, _P(C.c_double)]),
    "lucas_ctx_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "lucas_ctx_load": (C.c_int, [C.c_char_p, C.c_int, _P(C.c_void_p)]),
    "lucas_mersenne_test": (C.c_int, [C.c_int, C.c_int, _P(Options), _P(Result)]),
    "lucas_error_test": (C.c_int, [C.c_int, C.c_int, C.c_int, _P(Options), _P(Result)]),
    "lucas_autotune_N": (C.c_int, [C.c_int, C.c_int, _P(Options), _P(PlanInfo)]),
    "lucas_autotune_candidates": (C.c_int, [C.c_int, C.c_void_p, C.c_int, _P(C.c_int)]),
    "lucas_step_digits_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_longlong, C.c_int, C.c_void_p,
                                            _P(C.c_double)]),
    "lucas_batch_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, _P(C.c_void_p)]),
    "lucas_batch_destroy": (C.c_int, [C.c_void_p]),
    "lucas_batch_step": (C.c_int, [C.c_void_p, C.c_longlong]),
    "lucas_batch_res64": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "lucas_batch_get_digits": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "lucas_batch_time_iterations": (C.c_int, [C.c_void_p, C.c_longlong, _P(C.c_double)]),
    "lucas_server_start": (C.c_int, [_P(ServerOptions), _P(C.c_void_p)]),
    "lucas_server_submit": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_longlong, _P(C.c_longlong)]),
    "lucas_server_poll": (C.c_int, [C.c_void_p, C.c_longlong, _P(JobResult)]),
    "lucas_server_drain": (C.c_int, [C.c_void_p, C.c_double]),
    "lucas_server_stats": (C.c_int, [C.c_void_p, _P(C.c_longlong), _P(C.c_double), _P(C.c_int)]),
    "lucas_server_stop": (C.c_int, [C.c_void_p]),
}


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m gpulucas_amd.build` "
                              "(there is no CPU fallback)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def exported_symbols():
    return sorted(_SIGS)


def _check(rc: int, what: str):
    if rc != OK:
        raise LucasError(rc, what)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def version() -> str:
    return _lib().lucas_version().decode()


def device_count() -> int:
    n = C.c_int(0)
    _lib().lucas_device_count(C.byref(n))
    return n.value


def device_copy_bench(nbytes: int, reps: int = 20, device: int = 0) -> float:
    """Mean microseconds per device-to-device copy of `nbytes`."""
    us = C.c_double(0)
    _check(_lib().lucas_device_copy_bench(nbytes, reps, device, C.byref(us)), "device_copy_bench")
    return us.value


# ---- planning (host only) --------------------------------------------------

def default_N(p: int) -> int:
    return _lib().lucas_default_N(p)


def max_bits_per_word(N: int) -> float:
    """Calibrated bits-per-word line of the default length rule at length N."""
    return _lib().lucas_max_bits_per_word(N)


def plan(p: int, N: int = 0, A: int = 0, C_: int = 0) -> dict:
    info = PlanInfo()
    _check(_lib().lucas_plan_shape(p, N, A, C_, C.byref(info)), f"plan(p={p}, N={N})")
    return info.as_dict()


def plan_list(p: int, N: int) -> list[dict]:
    n = C.c_int(0)
    _check(_lib().lucas_plan_list(p, N, None, 0, C.byref(n)), f"plan_list(p={p}, N={N})")
    arr = (PlanInfo * n.value)()
    _check(_lib().lucas_plan_list(p, N, arr, n.value, C.byref(n)), "plan_list")
    return [a.as_dict() for a in arr]


def line_plan_list(p: int, N: int) -> list[dict]:
    """The line-kernel plans of (p, N): one wave per row or column, shapes
    compiled in (empty where none fits N and the carry depth)."""
    n = C.c_int(0)
    _check(_lib().lucas_line_plan_list(p, N, None, 0, C.byref(n)), f"line_plan_list(p={p}, N={N})")
    arr = (PlanInfo * max(1, n.value))()
    _check(_lib().lucas_line_plan_list(p, N, arr, n.value, C.byref(n)), "line_plan_list")
    return [arr[i].as_dict() for i in range(n.value)]


def autotune_shape(p: int, N: int = 0, iters: int = 200, max_candidates: int = 0, device: int = 0) -> tuple[dict, float]:
    info = PlanInfo()
    ms = C.c_double(0)
    opt = _options(device)
    _check(_lib().lucas_autotune_shape(p, N, iters, max_candidates, C.byref(opt), C.byref(info), C.byref(ms)),
           f"autotune_shape(p={p}, N={N})")
    return info.as_dict(), ms.value


def digit_tables(p: int, N: int):
    w = np.zeros(N, dtype=np.int32)
    a = np.zeros(N, dtype=np.float64)
    ai = np.zeros(N, dtype=np.float64)
    _check(_lib().lucas_digit_tables(p, N, _ptr(w), _ptr(a), _ptr(ai)), "digit_tables")
    return w, a, ai


def digits_res64(p: int, N: int, digits) -> tuple[int, bool, np.ndarray]:
    d = np.ascontiguousarray(digits, dtype=np.int32)
    out = np.zeros(N, dtype=np.int32)
    r = C.c_uint64(0)
    z = C.c_int(0)
    _check(_lib().lucas_digits_res64(p, N, _ptr(d), C.byref(r), C.byref(z), _ptr(out)), "digits_res64")
    return r.value, bool(z.value), out


# ---- single exponent ----------------------------------------------------------

class Context:
    """Device-resident state of one exponent (s_0 = 4 at creation)."""

    def __init__(self, p: int, N: int = 0, device: int = 0, A: int = 0, C_: int = 0, _handle=None,
                 plan: dict | None = None):
        self._h = C.c_void_p()
        if _handle is not None:
            self._h = _handle
        elif plan is not None:
            info = PlanInfo(**{k: plan[k] for k, _ in PlanInfo._fields_ if k in plan})
            _check(_lib().lucas_ctx_create_plan(C.byref(info), device, C.byref(self._h)), f"ctx_create_plan({plan})")
        else:
            _check(_lib().lucas_ctx_create_shape(p, N, A, C_, device, C.byref(self._h)),
                   f"ctx_create(p={p}, N={N})")
        info = PlanInfo()
        _check(_lib().lucas_ctx_info(self._h, C.byref(info)), "ctx_info")
        self.info = info.as_dict()
        self.p, self.N = self.info["p"], self.info["N"]

    @classmethod
    def load(cls, path: str, device: int = 0) -> "Context":
        h = C.c_void_p()
        _check(_lib().lucas_ctx_load(path.encode(), device, C.byref(h)), f"ctx_load({path})")
        return cls(0, _handle=h)

    def close(self):
        if self._h:
            _lib().lucas_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def iteration(self) -> int:
        return _lib().lucas_ctx_iteration(self._h)

    def step(self, iters: int):
        _check(_lib().lucas_ctx_step(self._h, iters), f"step({iters})")

    def digits(self) -> np.ndarray:
        d = np.zeros(self.N, dtype=np.int32)
        _check(_lib().lucas_ctx_get_digits(self._h, _ptr(d)), "get_digits")
        return d

    def device_digits(self) -> int:
        """Device address of the context's digit vector (N int32)."""
        ptr = C.c_void_p()
        _check(_lib().lucas_ctx_device_digits(self._h, C.byref(ptr)), "device_digits")
        return ptr.value

    def set_digits(self, digits, iteration: int = 0):
        d = np.ascontiguousarray(digits, dtype=np.int32)
        assert d.shape == (self.N,)
        _check(_lib().lucas_ctx_set_digits(self._h, _ptr(d), iteration), "set_digits")

    def res64(self) -> tuple[int, bool]:
        r = C.c_uint64(0)
        z = C.c_int(0)
        _check(_lib().lucas_ctx_res64(self._h, C.byref(r), C.byref(z)), "res64")
        return r.value, bool(z.value)

    def max_roundoff(self, reset: bool = False) -> float:
        v = C.c_double(0)
        _check(_lib().lucas_ctx_max_roundoff(self._h, int(reset), C.byref(v)), "max_roundoff")
        return v.value

    def product_bits(self, reset: bool = False) -> tuple[int, int]:
        """(largest bit length of a rounded product |c_j|, the reference's
        "max Bit Vector" = largest 64-bit leading-zero count of |c_j|) over
        every iteration since the last reset."""
        b, z = C.c_int(0), C.c_int(0)
        _check(_lib().lucas_ctx_product_bits(self._h, int(reset), C.byref(b), C.byref(z)), "product_bits")
        return b.value, z.value

    def prepare(self, iters: int):
        """Make a step of `iters` iterations launch-ready: build its replay
        graphs and run it once, restoring the state afterwards."""
        _check(_lib().lucas_ctx_prepare(self._h, iters), f"prepare({iters})")

    def time_kernels(self, iters: int = 100) -> tuple[float, float]:
        a, b = C.c_double(0), C.c_double(0)
        _check(_lib().lucas_ctx_time_kernels(self._h, iters, C.byref(a), C.byref(b)), "time_kernels")
        return a.value, b.value

    def phase_stamps(self, iters: int = 20, max_workgroups: int = 8192) -> tuple[np.ndarray, np.ndarray]:
        """Real-time stamps (100 MHz ticks, int64) of the last column and row
        launch after `iters` iterations: arrays of shape (workgroups, 6).  Needs
        the profiling library (libgpulucas_phases.so); the default library
        raises LucasError with status ERR_UNSUPPORTED."""
        col = np.zeros((max_workgroups, 8), np.uint64)
        row = np.zeros((max_workgroups, 8), np.uint64)
        nc, nr = C.c_int(0), C.c_int(0)
        _check(_lib().lucas_ctx_phase_stamps(self._h, iters, col.ctypes.data_as(_P(C.c_ulonglong)),
                                             row.ctypes.data_as(_P(C.c_ulonglong)), max_workgroups,
                                             C.byref(nc), C.byref(nr)), "phase_stamps")
        return col[:nc.value, :6].astype(np.int64), row[:nr.value, :6].astype(np.int64)

    def time_iterations(self, iters: int) -> float:
        v = C.c_double(0)
        _check(_lib().lucas_ctx_time_iterations(self._h, iters, C.byref(v)), "time_iterations")
        return v.value

    def save(self, path: str):
        _check(_lib().lucas_ctx_save(self._h, path.encode()), f"save({path})")


def _options(device=0, error_limit=0.0, checkpoint_every=0, checkpoint_path=None, resume=False):
    return Options(device, error_limit, checkpoint_every,
                   checkpoint_path.encode() if checkpoint_path else None, int(resume))


def mersenne_test(p: int, N: int = 0, device: int = 0, error_limit: float = 0.0,
                  checkpoint_every: int = 0, checkpoint_path: str | None = None, resume: bool = False) -> dict:
    """Full LL test; returns the result record and its status."""
    res = Result()
    opt = _options(device, error_limit, checkpoint_every, checkpoint_path, resume)
    rc = _lib().lucas_mersenne_test(p, N, C.byref(opt), C.byref(res))
    if rc not in (OK, ERR_ROUNDOFF):
        raise LucasError(rc, f"mersenne_test(p={p}, N={N})")
    d = res.as_dict()
    d["status"] = rc
    return d


def mersenneTest(p: int, N: int = 0, **kw) -> dict:  # noqa: N802 (reference name)
    return mersenne_test(p, N, **kw)


def error_test(p: int, N: int = 0, iterations: int = 0, device: int = 0, error_limit: float = 0.0) -> dict:
    res = Result()
    opt = _options(device, error_limit)
    rc = _lib().lucas_error_test(p, N, iterations, C.byref(opt), C.byref(res))
    if rc not in (OK, ERR_ROUNDOFF):
        raise LucasError(rc, f"error_test(p={p}, N={N})")
    d = res.as_dict()
    d["status"] = rc
    return d


def errorTest(p: int, N: int = 0, iterations: int = 0, **kw) -> float:  # noqa: N802
    """Mean ms/iteration, or -1.0 on a roundoff breach (errorTrial's contract)."""
    d = error_test(p, N, iterations, **kw)
    return -1.0 if d["status"] == ERR_ROUNDOFF else d["ms_per_iter"]


def autotune_candidates(p: int) -> list[int]:
    """Lengths the first window of autotune_N measures (host only)."""
    n = C.c_int(0)
    _check(_lib().lucas_autotune_candidates(p, None, 0, C.byref(n)), f"autotune_candidates({p})")
    arr = np.zeros(n.value, dtype=np.int32)
    _check(_lib().lucas_autotune_candidates(p, _ptr(arr), n.value, C.byref(n)), "autotune_candidates")
    return [int(x) for x in arr]


def autotune_N(p: int, iters: int = 300, device: int = 0, error_limit: float = 0.0) -> dict:
    info = PlanInfo()
    opt = _options(device, error_limit)
    _check(_lib().lucas_autotune_N(p, iters, C.byref(opt), C.byref(info)), f"autotune_N({p})")
    return info.as_dict()


def findSignalSize(p: int, **kw) -> int:  # noqa: N802
    return autotune_N(p, **kw)["N"]


# ---- no device allocation -----------------------------------------------------

NOALLOC_MAX_N = 524288


def step_digits_noalloc(p: int, N: int, digits_dev_ptr: int, iters: int, device: int = 0, stream: int = 0,
                        with_roundoff: bool = True):
    """Advance N int32 digits held in caller-owned device memory (address
    `digits_dev_ptr`) by `iters` iterations with no device allocation (the
    library's static per-GPU scratch array).  Returns the max roundoff (after
    synchronising the stream) or None."""
    err = C.c_double(0)
    _check(_lib().lucas_step_digits_noalloc(p, N, C.c_void_p(digits_dev_ptr), iters, device, C.c_void_p(stream),
                                            C.byref(err) if with_roundoff else None),
           f"step_digits_noalloc(p={p}, N={N})")
    return err.value if with_roundoff else None


# ---- batches ----------------------------------------------------------------

class Batch:
    """Many exponents advanced together (one launch per kernel per iteration)."""

    def __init__(self, ps, Ns=None, device: int = 0):
        ps = np.ascontiguousarray(ps, dtype=np.int32)
        self.n = len(ps)
        Ns_arr = np.ascontiguousarray(Ns if Ns is not None else np.zeros(self.n), dtype=np.int32)
        self._h = C.c_void_p()
        _check(_lib().lucas_batch_create(_ptr(ps), _ptr(Ns_arr), self.n, device, C.byref(self._h)),
               "batch_create")
        self.ps = ps
        self.Ns = [plan(int(p), int(n))["N"] for p, n in zip(ps, Ns_arr)]

    def close(self):
        if self._h:
            _lib().lucas_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, iters: int):
        _check(_lib().lucas_batch_step(self._h, iters), f"batch_step({iters})")

    def results(self):
        r = np.zeros(self.n, dtype=np.uint64)
        z = np.zeros(self.n, dtype=np.int32)
        e = np.zeros(self.n, dtype=np.float64)
        _check(_lib().lucas_batch_res64(self._h, _ptr(r), _ptr(z), _ptr(e)), "batch_res64")
        return [int(x) for x in r], [bool(x) for x in z], [float(x) for x in e]

    def digits(self, job: int) -> np.ndarray:
        d = np.zeros(self.Ns[job], dtype=np.int32)
        _check(_lib().lucas_batch_get_digits(self._h, job, _ptr(d)), "batch_get_digits")
        return d

    def time_iterations(self, iters: int) -> float:
        v = C.c_double(0)
        _check(_lib().lucas_batch_time_iterations(self._h, iters, C.byref(v)), "batch_time")
        return v.value


# ---- serving ------------------------------------------------------------------

class Server:
    """Continuous-batching request loop (device-resident state per job)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0, max_jobs: int = 0, chunk_iters: int = 0,
                 checkpoint_dir: str | None = None, checkpoint_every: int = 0, error_limit: float = 0.0):
        opt = ServerOptions(device, hbm_budget_bytes, max_jobs, chunk_iters,
                            checkpoint_dir.encode() if checkpoint_dir else None, checkpoint_every, error_limit)
        self._h = C.c_void_p()
        _check(_lib().lucas_server_start(C.byref(opt), C.byref(self._h)), "server_start")

    def submit(self, p: int, iterations: int, N: int = 0) -> int:
        jid = C.c_longlong(0)
        _check(_lib().lucas_server_submit(self._h, p, N, iterations, C.byref(jid)), f"submit({p})")
        return jid.value

    def poll(self, jid: int) -> dict:
        r = JobResult()
        _check(_lib().lucas_server_poll(self._h, jid, C.byref(r)), "poll")
        return r.as_dict()

    def drain(self, timeout_s: float = 0.0):
        _check(_lib().lucas_server_drain(self._h, timeout_s), "drain")

    def stats(self) -> dict:
        it, busy, peak = C.c_longlong(0), C.c_double(0), C.c_int(0)
        _check(_lib().lucas_server_stats(self._h, C.byref(it), C.byref(busy), C.byref(peak)), "stats")
        return {"iterations": it.value, "busy_s": busy.value, "peak_jobs": peak.value}

    def stop(self):
        if self._h:
            _lib().lucas_server_stop(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()
