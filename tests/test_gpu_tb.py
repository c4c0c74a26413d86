"""GPU parity of the temporally blocked kernels (T steps fused per launch).

In EXACT mode T fused steps round exactly like T single steps (also for
BF16/FP16 storage, whose intermediate levels are rounded to storage), so every
depth and tile must be bitwise equal to the oracle, including counts that are
not multiples of T (the remainder runs as plain steps: exactly `count` steps).
"""
import numpy as np
import pytest

import benchmarks_amd as d3
import oracle

pytestmark = pytest.mark.gpu

SHAPES = [(200, 90, 70), (600, 41, 33), (37, 130, 19)]


def run(ctx, f0, host_dtype, dtype, coef, count, **kw):
    nz, ny, nx = f0.shape
    with d3.Field(ctx, nx, ny, nz, dtype) as f:
        f.upload(f0, host_dtype)
        t = f.run(count, coef, **kw)
        return f.download(), t


@pytest.mark.parametrize("shape", SHAPES)
def test_tb_f32_every_tile_bitwise(ctx, shape):
    k = d3.constants_f32(*shape)
    f0 = d3.initialize_f32(*shape, k)
    f0 = d3.add_noise_f32(f0, 3, 1e-3)
    counts = (13, 8)
    want = {c: oracle.run_f32(f0, k[8:15], c) for c in counts}
    n = 0
    for t in d3.tiles():
        if t["dtype"] != d3.F32 or t["temporal_depth"] == 1 or t["arith"] != d3.ARITH_EXACT:
            continue
        if t["row_stream"]:
            continue                      # nz == 1 fields only (tests/test_gpu_2d.py)
        for c in counts:
            out, tm = run(ctx, f0, d3.F32, d3.F32, d3.coefficients(k), c,
                          depth=t["temporal_depth"], tile=t["index"])
            assert tm["steps"] == c
            assert tm["launches"] == c // t["temporal_depth"] + c % t["temporal_depth"]
            assert np.array_equal(out.view(np.uint32), want[c].view(np.uint32)), (t, c)
            n += 1
    assert n >= 10


def test_tb_f64_every_tile_bitwise(ctx):
    shape = (200, 90, 70)
    k = d3.constants_f64(*shape)
    f0 = d3.initialize_f64(*shape, k)
    want = oracle.run_f64(f0, k[8:15], 11)
    for t in d3.tiles():
        if t["dtype"] == d3.F64 and t["temporal_depth"] > 1 and t["arith"] == d3.ARITH_EXACT:
            out, _ = run(ctx, f0, d3.F64, d3.F64, d3.coefficients(k), 11,
                         depth=t["temporal_depth"], tile=t["index"])
            assert np.array_equal(out, want), t


@pytest.mark.parametrize("kind,dtype", [(0, d3.BF16), (1, d3.F16)])
def test_tb_low_precision_bitwise(ctx, kind, dtype):
    shape = (600, 41, 33)
    k = d3.constants_f32(*shape)
    f32 = d3.initialize_f32(*shape, k)
    want = oracle.run_lp(oracle.to_lp(f32, kind), k[8:15], 9, kind)
    for t in d3.tiles():
        if t["dtype"] == dtype and t["temporal_depth"] > 1 and t["arith"] == d3.ARITH_EXACT:
            out, _ = run(ctx, f32, d3.F32, dtype, d3.coefficients(k), 9,
                         depth=t["temporal_depth"], tile=t["index"])
            bits = out.view(np.uint16) if dtype == d3.F16 else out
            assert np.array_equal(bits, want), t


def test_tb_fma_within_tolerance(ctx):
    shape = (256, 256, 64)
    k = d3.constants_f32(*shape)
    f0 = d3.initialize_f32(*shape, k)
    want = oracle.run_f32(f0, k[8:15], 40)
    for depth in (2, 4, 8):
        out, _ = run(ctx, f0, d3.F32, d3.F32, d3.coefficients(k), 40, depth=depth,
                     arith=d3.ARITH_FMA)
        assert oracle.diff_norms(out, want)[0] <= 2e-6, depth
