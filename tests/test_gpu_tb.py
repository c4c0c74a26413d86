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
        if (t["dtype"] == d3.F64 and t["temporal_depth"] > 1 and t["arith"] == d3.ARITH_EXACT
                and not t["row_stream"]):         # 2D tiles: tests/test_gpu_2d.py
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
    """FMA mode (non-conforming: not the parity mode) at depths 2/4/8 within
    its own stated bound of 2e-6 relative L2 per 100 steps against the
    reference CPU path (here 40 steps)."""
    shape = (256, 256, 64)
    k = d3.constants_f32(*shape)
    f0 = d3.initialize_f32(*shape, k)
    want = oracle.run_f32(f0, k[8:15], 40)
    for depth in (2, 4, 8):
        out, _ = run(ctx, f0, d3.F32, d3.F32, d3.coefficients(k), 40, depth=depth,
                     arith=d3.ARITH_FMA)
        assert oracle.diff_norms(out, want)[0] <= 2e-6, depth


def test_autotuner_chooses_depth(tmp_path):
    """d3h_autotune with depth 0 searches depths 1/2/4 (3D fields) and
    remembers the fastest per size class; depth-0 runs and batches then use
    it, bitwise equal to plain steps, also after a save/load round trip.
    Depth 8 is not searched in 3D but runs when asked for."""
    shape = (160, 96, 72)
    k = d3.constants_f32(*shape)
    f0 = d3.add_noise_f32(d3.initialize_f32(*shape, k), 9, 1e-3)
    coef = d3.coefficients(k)
    want = oracle.run_f32(f0, k[8:15], 13)
    path = str(tmp_path / "tuning.txt")
    with d3.Context(0) as c:
        t = c.autotune(*shape, d3.F32, 0, d3.ARITH_EXACT)
        chosen = d3.tiles()[t["tile"]]
        print("chosen", chosen, t, flush=True)
        assert chosen["temporal_depth"] in (1, 2, 4)
        for T in (1, 2, 4, 8):          # every depth's result is remembered too
            out, tm = run(c, f0, d3.F32, d3.F32, coef, 13, depth=T)
            assert np.array_equal(out.view(np.uint32), want.view(np.uint32)), T
        out, tm = run(c, f0, d3.F32, d3.F32, coef, 13, depth=0)
        assert tm["tile"] == t["tile"] and tm["temporal_depth"] == chosen["temporal_depth"]
        assert np.array_equal(out.view(np.uint32), want.view(np.uint32))
        c.save_tuning(path)
        p = {"nx": shape[0], "ny": shape[1], "nz": shape[2], "dtype": d3.F32, "steps": 13,
             "coef": coef}
        with d3.Batch(c, [p, p]) as b:
            b.upload(0, f0, d3.F32)
            b.upload(1, f0, d3.F32)
            tb = b.run(depth=0, graph=False)
            assert np.array_equal(b.download(1, d3.F32).view(np.uint32), want.view(np.uint32))
    with open(path) as fh:
        lines = [l.split() for l in fh if not l.startswith("#")]
    assert any(l[1] == "0" and int(l[13]) == chosen["temporal_depth"] for l in lines), lines
    with d3.Context(0) as c2:
        out, tm = run(c2, f0, d3.F32, d3.F32, coef, 13, depth=0)
        assert tm["temporal_depth"] == 1          # nothing tuned: plain steps
        c2.load_tuning(path)
        out, tm = run(c2, f0, d3.F32, d3.F32, coef, 13, depth=0)
        assert tm["tile"] == t["tile"]
        assert np.array_equal(out.view(np.uint32), want.view(np.uint32))
