"""Plane schedules of depth-1 launches (d3h_run_config.pace_ns), through the C ABI.

A schedule only holds workgroups back before they request their next plane,
so every result must stay bitwise equal to the oracle with any schedule:
none, one every workgroup overtakes at once (1 ns per z iteration), one near
the tuned range and one far slower than the kernel (every wait then ends at
its bound).  The autotuner's schedule travels through the tune table.
"""
import numpy as np
import pytest

import benchmarks_amd as d3
import oracle

pytestmark = pytest.mark.gpu

PACES = (-1, 1, 700, 3000)


def noisy_f32(shape, seed=5):
    k = d3.constants_f32(*shape)
    f0 = d3.initialize_f32(*shape, k)
    rng = np.random.default_rng(seed)
    return k, (f0 + rng.standard_normal(f0.shape).astype(np.float32) * np.float32(1e-3))


def run(ctx, f0, coef, count, dtype=d3.F32, yrev=False, **kw):
    nz, ny, nx = f0.shape
    with d3.Field(ctx, nx, ny, nz, dtype) as f:
        if yrev:
            f.set_y_reversed(True)
        f.upload(f0)
        t = f.run(count, coef, **kw)
        return f.download(), t


@pytest.mark.parametrize("shape", [(200, 37, 61), (512, 9, 40)])
def test_every_plain_3d_tile_with_schedules_matches_oracle(ctx, shape):
    k, f0 = noisy_f32(shape)
    want = oracle.run_f32(f0, k[8:15], 5)
    coef = d3.coefficients(k)
    for t in d3.tiles():
        if (t["dtype"] != d3.F32 or t["temporal_depth"] != 1 or t["arith"] != d3.ARITH_EXACT
                or t["row_stream"]):
            continue
        if t["tile_y"] <= 2 and shape[1] > t["tile_y"]:
            continue
        for pace in PACES:
            for zc in (0, 3):
                print("tile", t["index"], shape, "pace", pace, "zchunks", zc, flush=True)
                out, tm = run(ctx, f0, coef, 5, tile=t["index"], zchunks=zc, pace_ns=pace)
                assert tm["pace_ns"] == max(pace, 0), tm
                assert np.array_equal(out.view(np.uint32), want.view(np.uint32)), (t, pace, zc)


def test_every_f64_plain_tile_with_schedules_matches_oracle(ctx):
    shape = (200, 37, 41)
    k = d3.constants_f64(*shape)
    f0 = d3.initialize_f64(*shape, k)
    want = oracle.run_f64(f0, k[8:15], 5)
    tiles = [t for t in d3.tiles() if t["dtype"] == d3.F64 and t["temporal_depth"] == 1
             and t["arith"] == d3.ARITH_EXACT and not t["row_stream"]]
    assert tiles
    for t in tiles:
        for pace in PACES:
            print("F64 tile", t["index"], "pace", pace, flush=True)
            out, tm = run(ctx, f0, d3.coefficients(k), 5, dtype=d3.F64, tile=t["index"], zchunks=3,
                          pace_ns=pace)
            assert tm["pace_ns"] == max(pace, 0), tm
            assert np.array_equal(out.view(np.uint64), want.view(np.uint64)), (t, pace)


def test_schedules_beyond_a_millisecond_are_refused(ctx, tmp_path):
    with d3.Field(ctx, 16, 16, 16) as f:
        f.upload(np.zeros((16, 16, 16), np.float32))
        with pytest.raises(d3.D3HError):
            f.run(1, [0.1] * 6 + [0.4], pace_ns=1000001)
        tm = f.run(1, [0.1] * 6 + [0.4], pace_ns=1000000)
        assert tm["pace_ns"] in (0, 1000000)
    bad = tmp_path / "slow.txt"
    # dtype depth fma log2nx log2ny log2nz vec G RPW NW pf zchunks ms depth batch rows pace_ns
    bad.write_text("0 1 0 4 4 4 4 2 1 4 1 0 0.1 1 0 0 1000001\n")
    with pytest.raises(d3.D3HError):
        ctx.load_tuning(str(bad))


def test_fused_depths_ignore_schedules(ctx):
    shape = (130, 20, 33)
    k, f0 = noisy_f32(shape)
    want = oracle.run_f32(f0, k[8:15], 9)
    for depth in (2, 4):
        out, tm = run(ctx, f0, d3.coefficients(k), 9, depth=depth, pace_ns=700)
        assert tm["temporal_depth"] == depth and tm["pace_ns"] == 0, tm
        assert np.array_equal(out.view(np.uint32), want.view(np.uint32)), depth


def test_2d_and_16_bit_tiles_with_schedules(ctx):
    nx, ny = 700, 90
    k2 = d3.constants2d_f32(nx, ny)
    f2 = d3.add_noise_f32(d3.initialize2d_f32(nx, ny, k2).reshape(-1), 3, 1e-3).reshape(ny, nx)
    want2 = oracle.run2d_f32(f2, k2[8:15], 7)
    c2 = d3.coefficients(k2)
    c2[4] = c2[5] = 0.0
    rows = [t for t in d3.tiles() if t["row_stream"] and t["dtype"] == d3.F32
            and t["temporal_depth"] == 1 and t["arith"] == d3.ARITH_EXACT]
    assert rows
    for t in rows:
        for pace in (1, 700, 3000):
            print("2D tile", t["index"], "pace", pace, flush=True)
            out, tm = run(ctx, f2.reshape(1, ny, nx), c2, 7, yrev=True, tile=t["index"], zchunks=3,
                          pace_ns=pace)
            assert tm["pace_ns"] == pace
            assert np.array_equal(out.reshape(ny, nx).view(np.uint32), want2.view(np.uint32)), (t, pace)
    # 16-bit storage with stochastic rounding
    shape = (96, 40, 30)
    k, f0 = noisy_f32(shape)
    bits = oracle.to_lp(f0, 0)
    coef = np.float32(d3.coefficients(k))
    want = oracle.run_bf16_sr(bits, coef, 6, 77)
    for pace in (-1, 700):
        with d3.Field(ctx, *shape, d3.BF16) as f:
            f.upload(bits, d3.BF16)
            tm = f.run(6, coef, rounding=d3.ROUND_STOCHASTIC, seed=77, pace_ns=pace)
            got = f.download().view(np.uint16)
        assert tm["pace_ns"] == max(pace, 0)
        assert np.array_equal(got, want), pace


def test_tuned_schedule_travels_through_the_tune_table(tmp_path):
    shape = (256, 256, 128)
    k, f0 = noisy_f32(shape)
    want = oracle.run_f32(f0, k[8:15], 4)
    coef = d3.coefficients(k)
    path = str(tmp_path / "tune.txt")
    with d3.Context(0) as c:
        tuned = c.autotune(*shape, d3.F32, 1)
        out, tm = run(c, f0, coef, 4)
        print("tuned", tuned, tm, flush=True)
        assert tm["tile"] == tuned["tile"] and tm["pace_ns"] >= 0
        assert np.array_equal(out.view(np.uint32), want.view(np.uint32))
        pace = tm["pace_ns"]
        # an explicit "none" overrides the tuned schedule, same bits
        out, tm2 = run(c, f0, coef, 4, pace_ns=-1)
        assert tm2["pace_ns"] == 0 and tm2["tile"] == tuned["tile"]
        assert np.array_equal(out.view(np.uint32), want.view(np.uint32))
        c.save_tuning(path)
    with open(path) as fh:
        lines = [l.split() for l in fh if not l.startswith("#")]
    assert lines and all(len(l) == 17 for l in lines), lines
    with d3.Context(0) as c2:
        c2.load_tuning(path)
        out, tm = run(c2, f0, coef, 4)
        assert tm["tile"] == tuned["tile"] and tm["pace_ns"] == pace, (tm, pace)
        assert np.array_equal(out.view(np.uint32), want.view(np.uint32))
    # a table written before schedules (16 columns) loads with none
    old = tmp_path / "old.txt"
    old.write_text("\n".join(" ".join(l[:16]) for l in lines) + "\n")
    with d3.Context(0) as c3:
        c3.load_tuning(str(old))
        out, tm = run(c3, f0, coef, 4)
        assert tm["tile"] == tuned["tile"] and tm["pace_ns"] == 0
        assert np.array_equal(out.view(np.uint32), want.view(np.uint32))
    # a negative schedule in a table is refused
    bad = tmp_path / "bad.txt"
    bad.write_text(" ".join(lines[0][:16] + ["-5"]) + "\n")
    with d3.Context(0) as c4:
        with pytest.raises(d3.D3HError):
            c4.load_tuning(str(bad))


def test_full_size_tuned_schedule_equals_unscheduled():
    """512^3 FP32 (the benchmark's grid), the autotuner's tile, z-chunks and
    schedule against the same tile unscheduled: bitwise equal after 20 steps
    (device-side comparison)."""
    n = 512
    k = d3.constants_f32(n, n, n)
    f0 = d3.initialize_f32(n, n, n, k)
    coef = d3.coefficients(k)
    with d3.Context(0) as c:
        tuned = c.autotune(n, n, n, d3.F32, 1)
        with d3.Field(c, n, n, n) as a, d3.Field(c, n, n, n) as b:
            a.upload(f0)
            b.upload(f0)
            ta = a.run(20, coef)
            tb = b.run(20, coef, tile=ta["tile"], zchunks=ta["zchunks"], pace_ns=-1)
            print("tuned", tuned, "scheduled", ta, "unscheduled", tb, flush=True)
            assert ta["tile"] == tuned["tile"] and tb["pace_ns"] == 0
            assert a.compare(b)["max_abs"] == 0.0


def test_scheduled_graph_replays_match_oracle(ctx):
    """A scheduled depth-1 run captured in a graph and replayed (the schedule
    is part of the graph's key): every replay bitwise equal to the oracle."""
    shape = (256, 40, 70)
    k, f0 = noisy_f32(shape)
    coef = d3.coefficients(k)
    want = {c: oracle.run_f32(f0, k[8:15], c) for c in (6, 12, 18)}
    tile = [t for t in d3.tiles() if t["dtype"] == d3.F32 and t["temporal_depth"] == 1
            and t["arith"] == d3.ARITH_EXACT and not t["row_stream"] and t["tile_y"] > 2][0]
    nz, ny, nx = f0.shape
    with d3.Field(ctx, nx, ny, nz) as f:
        f.upload(f0)
        for c in (6, 12, 18):
            tm = f.run(6, coef, tile=tile["index"], zchunks=3, pace_ns=700, graph=True)
            assert tm["pace_ns"] == 700
            out = f.download()
            assert np.array_equal(out.view(np.uint32), want[c].view(np.uint32)), c
        # a different schedule re-captures; still the same bits
        f.upload(f0)
        f.run(6, coef, tile=tile["index"], zchunks=3, pace_ns=1500, graph=True)
        assert np.array_equal(f.download().view(np.uint32), want[6].view(np.uint32))
