"""Seeded synthetic workloads of the batched and served configurations.

* batched: `count` independent problems, each dimension floor(exp(U(ln lo,
  ln hi))) drawn independently (seeded), kappa ~ U[0.05, 0.2], a fixed step
  count, one storage dtype.  Coefficients use dt = 0.1 h_min^2 / kappa (see
  engine.stable_coefficients), so kappa cancels and the coefficients depend on
  the shape only.
* served: requests with base edge floor(exp(U(ln lo, ln hi))), slab aspect
  factors exp(U(ln 1/2, ln 2)) on y and z (clipped to [lo/2, hi]), steps
  ~ U[s_lo, s_hi], 70 % FP32 / 30 % BF16 storage, Poisson arrival times at
  rate lam (seeded).
Initial fields are the cosine product of each problem's own grid plus seeded
noise of amplitude 1e-3 (counter-based, identical on every host).
"""
import ctypes

import numpy as np

from . import _native as N
from .engine import add_noise_f32, constants_f32, initialize_f32, stable_coefficients


def batched_problems(count=512, lo=32, hi=256, steps=200, dtype=N.F32, seed=4,
                     rounding=N.ROUND_NEAREST):
    """rounding applies to BF16 problems (ROUND_STOCHASTIC: seeded per problem)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nx, ny, nz = [int(np.exp(rng.uniform(np.log(lo), np.log(hi)))) for _ in range(3)]
        kappa = float(rng.uniform(0.05, 0.2))
        coef, dt = stable_coefficients(nx, ny, nz, kappa)
        p = {"nx": nx, "ny": ny, "nz": nz, "dtype": dtype, "steps": steps,
             "coef": coef, "kappa": kappa, "dt": dt}
        if dtype == N.BF16 and rounding == N.ROUND_STOCHASTIC:
            p.update(rounding=rounding, seed=0x5EED0000 + i)
        out.append(p)
    return out


def served_requests(count, lam, lo=32, hi=384, s_lo=50, s_hi=2000, seed=5,
                    rounding=N.ROUND_NEAREST):
    """rounding applies to the BF16 requests (ROUND_STOCHASTIC: seeded per request)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += float(rng.exponential(1.0 / lam))
        e = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        ay, az = np.exp(rng.uniform(np.log(0.5), np.log(2.0), 2))
        nx = e
        ny = int(np.clip(round(e * ay), lo // 2, hi))
        nz = int(np.clip(round(e * az), lo // 2, hi))
        steps = int(rng.integers(s_lo, s_hi + 1))
        dtype = N.BF16 if rng.random() < 0.3 else N.F32
        coef, dt = stable_coefficients(nx, ny, nz)
        p = {"nx": nx, "ny": ny, "nz": nz, "dtype": dtype, "steps": steps, "coef": coef,
             "dt": dt, "arrival_s": t}
        if dtype == N.BF16 and rounding == N.ROUND_STOCHASTIC:
            p.update(rounding=rounding, seed=0x5EED0000 + i)
        out.append(p)
    return out


def initial_field(p, index, out=None):
    k = constants_f32(p["nx"], p["ny"], p["nz"])
    f = initialize_f32(p["nx"], p["ny"], p["nz"], k, 0.0, out=out)
    return add_noise_f32(f, 1000003 + index, 1e-3)


class PinnedArena:
    """Pinned host arrays (hipHostMalloc) for request inputs/outputs, so the
    server's copies run asynchronously on its copy stream."""

    def __init__(self):
        self.ptrs = []

    def empty(self, shape, dtype=np.float32):
        n = int(np.prod(shape))
        nbytes = max(1, n * np.dtype(dtype).itemsize)
        p = ctypes.c_void_p()
        N.check(N.lib().d3h_host_alloc(nbytes, ctypes.byref(p)), "d3h_host_alloc")
        self.ptrs.append(p)
        buf = (ctypes.c_byte * nbytes).from_address(p.value)
        return np.frombuffer(buf, dtype=dtype, count=n).reshape(shape)

    def close(self):
        for p in self.ptrs:
            N.lib().d3h_host_free(p)
        self.ptrs = []
