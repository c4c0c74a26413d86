"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this,
as the checker / CPU baseline.  The product (benchmarks_amd) never imports it.
Arrays are (nz, ny, nx) numpy arrays, x fastest.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")

_f = ctypes.POINTER(ctypes.c_float)
_d = ctypes.POINTER(ctypes.c_double)
_u16 = ctypes.POINTER(ctypes.c_uint16)
_lib = None


def build():
    """Compile liboracle.so (gcc, seconds)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        sig = {
            "oracle_constants_f32": [ctypes.c_int] * 3 + [_f],
            "oracle_constants_f64": [ctypes.c_int] * 3 + [_d],
            "oracle_initialize_f32": [_f, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f, ctypes.c_float],
            "oracle_initialize_f64": [_d, ctypes.c_int, ctypes.c_int, ctypes.c_int, _d, ctypes.c_double],
            "oracle_step_f32": [_f, _f, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f],
            "oracle_step_f64": [_d, _d, ctypes.c_int, ctypes.c_int, ctypes.c_int, _d],
            "oracle_run_f32": [_f, _f, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f, ctypes.c_int],
            "oracle_run_f64": [_d, _d, ctypes.c_int, ctypes.c_int, ctypes.c_int, _d, ctypes.c_int],
            "oracle_run_lp": [_u16, _u16, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f, ctypes.c_int,
                              ctypes.c_int],
            "oracle_f32_to_lp": [_f, _u16, ctypes.c_long, ctypes.c_int],
            "oracle_run_bf16_sr": [_u16, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f, ctypes.c_int,
                                   ctypes.c_uint32, ctypes.c_int],
            "oracle_sr_hash": [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32],
            "oracle_sr_round": [ctypes.c_float, ctypes.c_uint32, ctypes.c_int],
            "oracle_run_lp_sr": [_u16, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_uint32, ctypes.c_int],
            "oracle_lp_to_f32": [_u16, _f, ctypes.c_long, ctypes.c_int],
            "oracle_accuracy_f32": [_f, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f, ctypes.c_int],
            "oracle_rms_vs_analytical_f32": [_f, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f,
                                             ctypes.c_int],
            "oracle_rms_vs_analytical_f64": [_d, ctypes.c_int, ctypes.c_int, ctypes.c_int, _d,
                                             ctypes.c_int],
            "oracle_add_noise_f32": [_f, ctypes.c_long, ctypes.c_uint64, ctypes.c_float],
            "oracle_constants2d_f32": [ctypes.c_int, ctypes.c_int, _f],
            "oracle_initialize2d_f32": [_f, ctypes.c_int, ctypes.c_int, _f, ctypes.c_float],
            "oracle_run2d_f32": [_f, _f, ctypes.c_int, ctypes.c_int, _f, ctypes.c_int],
            "oracle_accuracy2d_f32": [_f, ctypes.c_int, ctypes.c_int, _f, ctypes.c_int],
            "oracle_diff_norms_f32": [_f, _f, ctypes.c_long, _d],
            "oracle_diff_norms_f64": [_d, _d, ctypes.c_long, _d],
        }
        res = {"oracle_sr_hash": ctypes.c_uint32, "oracle_sr_round": ctypes.c_uint16, "oracle_accuracy_f32": ctypes.c_float, "oracle_accuracy2d_f32": ctypes.c_float, "oracle_rms_vs_analytical_f32": ctypes.c_double,
               "oracle_rms_vs_analytical_f64": ctypes.c_double}
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res.get(name, None)
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def constants_f32(nx, ny, nz):
    o = np.zeros(15, np.float32)
    lib().oracle_constants_f32(nx, ny, nz, _p(o, _f))
    return o


def constants_f64(nx, ny, nz):
    o = np.zeros(15, np.float64)
    lib().oracle_constants_f64(nx, ny, nz, _p(o, _d))
    return o


def initialize_f32(nx, ny, nz, k, time=0.0):
    out = np.empty((nz, ny, nx), np.float32)
    k = np.ascontiguousarray(k, np.float32)
    lib().oracle_initialize_f32(_p(out, _f), nx, ny, nz, _p(k, _f), float(time))
    return out


def initialize_f64(nx, ny, nz, k, time=0.0):
    out = np.empty((nz, ny, nx), np.float64)
    k = np.ascontiguousarray(k, np.float64)
    lib().oracle_initialize_f64(_p(out, _d), nx, ny, nz, _p(k, _d), float(time))
    return out


def run_f32(f, coef, count):
    """`count` steps of a float32 field (returns a new array)."""
    a = np.array(f, dtype=np.float32, copy=True, order="C")
    b = np.empty_like(a)
    c = np.ascontiguousarray(coef, np.float32)
    nz, ny, nx = a.shape
    lib().oracle_run_f32(_p(a, _f), _p(b, _f), nx, ny, nz, _p(c, _f), count)
    return a


def run_f64(f, coef, count):
    a = np.array(f, dtype=np.float64, copy=True, order="C")
    b = np.empty_like(a)
    c = np.ascontiguousarray(coef, np.float64)
    nz, ny, nx = a.shape
    lib().oracle_run_f64(_p(a, _d), _p(b, _d), nx, ny, nz, _p(c, _d), count)
    return a


def step_f32(f, coef):
    a = np.ascontiguousarray(f, np.float32)
    b = np.empty_like(a)
    c = np.ascontiguousarray(coef, np.float32)
    nz, ny, nx = a.shape
    lib().oracle_step_f32(_p(a, _f), _p(b, _f), nx, ny, nz, _p(c, _f))
    return b


def to_lp(f32, kind):
    """float32 -> uint16 bits (kind 0 = bfloat16, 1 = binary16)."""
    a = np.ascontiguousarray(f32, np.float32)
    out = np.empty(a.shape, np.uint16)
    lib().oracle_f32_to_lp(_p(a, _f), _p(out, _u16), a.size, kind)
    return out


def from_lp(bits, kind):
    a = np.ascontiguousarray(bits, np.uint16)
    out = np.empty(a.shape, np.float32)
    lib().oracle_lp_to_f32(_p(a, _u16), _p(out, _f), a.size, kind)
    return out


def run_lp(bits, coef, count, kind):
    a = np.array(bits, dtype=np.uint16, copy=True, order="C")
    b = np.empty_like(a)
    c = np.ascontiguousarray(coef, np.float32)
    nz, ny, nx = a.shape
    lib().oracle_run_lp(_p(a, _u16), _p(b, _u16), nx, ny, nz, _p(c, _f), count, kind)
    return a


def run_bf16_sr(bits, coef, count, seed, step0=0):
    """`count` steps of BF16 storage with stochastic rounding (new array)."""
    a = np.array(bits, dtype=np.uint16, copy=True, order="C")
    c = np.ascontiguousarray(coef, np.float32)
    nz, ny, nx = a.shape
    lib().oracle_run_bf16_sr(_p(a, _u16), nx, ny, nz, _p(c, _f), count, seed, step0)
    return a


def run_lp_sr(bits, coef, count, kind, seed, step0=0):
    """`count` steps of 16-bit storage (kind 0 BF16, 1 FP16) with stochastic
    rounding (new array)."""
    a = np.array(bits, dtype=np.uint16, copy=True, order="C")
    c = np.ascontiguousarray(coef, np.float32)
    nz, ny, nx = a.shape
    lib().oracle_run_lp_sr(_p(a, _u16), nx, ny, nz, _p(c, _f), count, kind, seed, step0)
    return a


def sr_hash(seed, row, pair, step):
    """The 32 random bits of x pair `pair` of row `row` (= z*ny + y) at `step`."""
    return int(lib().oracle_sr_hash(seed, row, pair, step))


def sr_round(value, r, kind):
    """One float32 value stochastically rounded to 16 bits with random bits r
    (kind 0 BF16, 1 FP16): the bit pattern."""
    return int(lib().oracle_sr_round(float(value), r, kind))


def accuracy_f32(f, k, count):
    nz, ny, nx = f.shape
    f = np.ascontiguousarray(f, np.float32)
    k = np.ascontiguousarray(k, np.float32)
    return float(lib().oracle_accuracy_f32(_p(f, _f), nx, ny, nz, _p(k, _f), count))


def rms_vs_analytical_f32(f, k, count):
    nz, ny, nx = f.shape
    f = np.ascontiguousarray(f, np.float32)
    k = np.ascontiguousarray(k, np.float32)
    return float(lib().oracle_rms_vs_analytical_f32(_p(f, _f), nx, ny, nz, _p(k, _f), count))


def rms_vs_analytical_f64(f, k, count):
    nz, ny, nx = f.shape
    f = np.ascontiguousarray(f, np.float64)
    k = np.ascontiguousarray(k, np.float64)
    return float(lib().oracle_rms_vs_analytical_f64(_p(f, _d), nx, ny, nz, _p(k, _d), count))


def constants2d_f32(nx, ny):
    o = np.zeros(15, np.float32)
    lib().oracle_constants2d_f32(nx, ny, _p(o, _f))
    return o


def initialize2d_f32(nx, ny, k, time=0.0):
    out = np.empty((ny, nx), np.float32)
    k = np.ascontiguousarray(k, np.float32)
    lib().oracle_initialize2d_f32(_p(out, _f), nx, ny, _p(k, _f), float(time))
    return out


def run2d_f32(f, coef, count):
    """`count` 5-point steps of a float32 (ny, nx) field (returns a new array);
    coef in the 7-slot order, ct and cb ignored."""
    a = np.array(f, dtype=np.float32, copy=True, order="C")
    b = np.empty_like(a)
    c = np.ascontiguousarray(coef, np.float32)
    ny, nx = a.shape
    lib().oracle_run2d_f32(_p(a, _f), _p(b, _f), nx, ny, _p(c, _f), count)
    return a


def accuracy2d_f32(f, k, count):
    ny, nx = f.shape
    f = np.ascontiguousarray(f, np.float32)
    k = np.ascontiguousarray(k, np.float32)
    return float(lib().oracle_accuracy2d_f32(_p(f, _f), nx, ny, _p(k, _f), count))


def add_noise_f32(f, seed, amp):
    a = np.array(f, dtype=np.float32, copy=True, order="C")
    lib().oracle_add_noise_f32(_p(a, _f), a.size, seed, float(amp))
    return a


def diff_norms(a, b):
    """(relative L2, max abs, L2) of a - b, FP64."""
    out = np.zeros(3, np.float64)
    if a.dtype == np.float64:
        lib().oracle_diff_norms_f64(_p(np.ascontiguousarray(a), _d), _p(np.ascontiguousarray(b, np.float64), _d),
                                    a.size, _p(out, _d))
    else:
        lib().oracle_diff_norms_f32(_p(np.ascontiguousarray(a, np.float32), _f),
                                    _p(np.ascontiguousarray(b, np.float32), _f), a.size, _p(out, _d))
    return tuple(out)
