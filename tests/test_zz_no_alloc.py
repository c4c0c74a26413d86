"""The allocation-free entry point d3h_field_error_analytical_noalloc against
d3h_field_error_analytical, through the C ABI: two calls in a row on one
device (the second on a smaller field of another dtype, so the static
scratch keeps the first call's contents), and a third after the allocating
entry.  Bar: the norms agree to 1e-13 absolute (the FP64 tolerance on O(1)
field values; the two entries differ only in the order of the FP64 sums and
in evaluating cos on the device instead of the host).  Sizes beyond the
scratch are refused with D3H_ERR_INVALID."""
import ctypes

import pytest

import benchmarks_amd as d3
from benchmarks_amd import _native as N

pytestmark = pytest.mark.gpu

TOL = 1e-13


def _norms(fn, field, kappa, time, k, d):
    out = (ctypes.c_double * 2)()
    rc = fn(field.handle, kappa, time, (ctypes.c_double * 3)(*k), (ctypes.c_double * 3)(*d), out)
    assert rc == N.D3H_OK, N.lib().d3h_last_error()
    return out[0], out[1]


def _check(L, field, kappa, time, k, d):
    got = _norms(L.d3h_field_error_analytical_noalloc, field, kappa, time, k, d)
    want = _norms(L.d3h_field_error_analytical, field, kappa, time, k, d)
    print("noalloc", got, "allocating", want, flush=True)
    assert want[0] > 0 and want[1] > 0
    assert abs(got[0] - want[0]) <= TOL and abs(got[1] - want[1]) <= TOL
    return got


def test_noalloc_error_norms_equal_the_allocating_entry():
    L = N.lib()
    cases = [((300, 170, 90), d3.F32, 7), ((41, 29, 17), d3.BF16, 5), ((64, 48, 40), d3.F64, 3)]
    with d3.Context(0) as ctx:
        for shape, dtype, steps in cases:
            if dtype == d3.F64:
                k = d3.constants_f64(*shape)
                f0 = d3.initialize_f64(*shape, k)
            else:
                k = d3.constants_f32(*shape)
                f0 = d3.add_noise_f32(d3.initialize_f32(*shape, k), 13, 1e-4)
            with d3.Field(ctx, *shape, dtype) as f:
                f.upload(f0, d3.F64 if dtype == d3.F64 else d3.F32)
                f.run(steps, d3.coefficients(k))
                t = float(steps) * float(k[7])
                a = _check(L, f, float(k[0]), t, [float(v) for v in k[4:7]],
                           [float(v) for v in k[1:4]])
                # the same call again gives the same bits (the scratch is refilled)
                b = _norms(L.d3h_field_error_analytical_noalloc, f, float(k[0]), t,
                           [float(v) for v in k[4:7]], [float(v) for v in k[1:4]])
                assert a == b
        # beyond the static factor table: refused, nothing launched
        with d3.Field(ctx, 65000, 600, 1) as f:
            out = (ctypes.c_double * 2)()
            rc = L.d3h_field_error_analytical_noalloc(f.handle, 0.1, 0.0,
                                                      (ctypes.c_double * 3)(1.0, 1.0, 1.0),
                                                      (ctypes.c_double * 3)(1.0, 1.0, 1.0), out)
            assert rc == N.D3H_ERR_INVALID
            assert b"exceeds" in L.d3h_last_error()
