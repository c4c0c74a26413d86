"""GPU parity of the diffusion/ family on the HIP path, through the C ABI.

The 2D 5-point operator (diffusion/diffusion_openmp.cc:59-82) runs as an
nx x ny x 1 device field with reversed rows and ct = cb = 0; the family's 3D
form (y-1 summed before y+1) runs as a reversed 3D field.  EXACT mode must be
bitwise equal to the family's own CPU build (fixtures from
tests/golden/gen_golden2d.py) and to the oracle.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import benchmarks_amd as d3
import oracle
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def golden2d():
    with open(os.path.join(GOLDEN, "reference_cpu_2d.json")) as f:
        meta = json.load(f)
    return meta, dict(np.load(os.path.join(GOLDEN, "reference_cpu_2d.npz")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def check_case(out, case, arrays, prefix):
    key = "x".join(map(str, case["shape"])) + "_c%d" % case["count"]
    flat = np.ascontiguousarray(out).reshape(-1)
    if prefix + key in arrays:
        assert np.array_equal(flat.view(np.uint32), arrays[prefix + key].view(np.uint32)), key
    else:
        idx = np.arange(0, flat.size, case["sample_stride"])
        assert list(flat[idx].view(np.uint32)) == case["samples_bits"], key
    assert sha(flat) == case["sha256"], key


def run2d(ctx, f0, coef, count, **kw):
    ny, nx = f0.shape
    c = [float(v) for v in coef]
    c[4] = c[5] = 0.0
    with d3.Field(ctx, nx, ny, 1) as f:
        f.set_y_reversed(True)
        f.upload(f0)
        f.run(count, c, **kw)
        return f.download().reshape(ny, nx)


def test_reversed_upload_download_roundtrip(ctx):
    rng = np.random.default_rng(11)
    for dtype, host in ((d3.F32, rng.random((5, 37, 29), dtype=np.float32)),
                        (d3.F64, rng.random((3, 8, 70))),
                        (d3.BF16, rng.integers(0x3c00, 0x4200, (4, 9, 33)).astype(np.uint16))):
        nz, ny, nx = host.shape
        with d3.Field(ctx, nx, ny, nz, dtype) as f:
            f.set_y_reversed(True)
            f.upload(host)
            back = f.download()
            assert back.dtype == host.dtype
            assert np.array_equal(back.view(np.uint8), host.view(np.uint8)), dtype
            # the device holds reversed rows: a plain field built from the
            # reversed host array compares equal to it
            with d3.Field(ctx, nx, ny, nz, dtype) as g:
                g.upload(np.ascontiguousarray(host[:, ::-1, :]))
                assert f.compare(g)["max_abs"] == 0.0


def test_2d_reference_fixtures_bitwise(ctx, golden2d):
    meta, arrays = golden2d
    for case in meta["cases2d"]:
        nx, ny = case["shape"]
        print("2d case", case["shape"], case["count"], flush=True)
        k = d3.constants2d_f32(nx, ny)
        out = run2d(ctx, d3.initialize2d_f32(nx, ny, k), d3.coefficients(k), case["count"])
        check_case(out, case, arrays, "out2d_")
        acc = oracle.accuracy2d_f32(out, k, case["count"])
        assert int(np.float32(acc).view(np.uint32)) == case["accuracy_bits"]


@pytest.mark.parametrize("depth", [2, 4])
def test_2d_temporal_blocking_bitwise(ctx, golden2d, depth):
    meta, arrays = golden2d
    for case in meta["cases2d"]:
        if case["count"] not in (7, 100):
            continue
        nx, ny = case["shape"]
        k = d3.constants2d_f32(nx, ny)
        out = run2d(ctx, d3.initialize2d_f32(nx, ny, k), d3.coefficients(k), case["count"],
                    depth=depth)
        check_case(out, case, arrays, "out2d_")


def test_2d_large_matches_oracle(ctx):
    nx, ny = 4096, 2048
    k = d3.constants2d_f32(nx, ny)
    f0 = d3.initialize2d_f32(nx, ny, k)
    f0 = d3.add_noise_f32(f0.reshape(-1), 7, 1e-3).reshape(ny, nx)
    out = run2d(ctx, f0, d3.coefficients(k), 20)
    ref = oracle.run2d_f32(f0, k[8:15], 20)
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


def test_run_kernel2d_entry(ctx, golden2d):
    meta, arrays = golden2d
    k = d3.constants2d_f32(33, 17)
    f = d3.initialize2d_f32(33, 17, k)
    d3.run_kernel2d(7, f, d3.coefficients(k))
    assert np.array_equal(f.reshape(-1).view(np.uint32), arrays["out2d_33x17_c7"].view(np.uint32))


def test_family_3d_fixtures_bitwise(ctx, golden2d):
    meta, arrays = golden2d
    for case in meta["cases3d"]:
        nx, ny, nz = case["shape"]
        c = np.array(case["constants_bits"], np.uint32).view(np.float32)
        f0 = d3.initialize_f32(nx, ny, nz, c)
        with d3.Field(ctx, nx, ny, nz) as f:
            f.set_y_reversed(True)
            f.upload(f0)
            f.run(case["count"], d3.coefficients(c))
            out = f.download()
        check_case(out, case, arrays, "out3d_")


def test_driver_nd2_matches_family_output(golden2d, tmp_path):
    """`diffusion3d_hip --nd 2` prints the diffusion/ family's lines, its
    accuracy for 256x256 / 100 steps, and dumps the family's "%.7E" file."""
    import subprocess
    from benchmarks_amd import _native as N
    meta, _ = golden2d
    case = [c for c in meta["cases2d"] if c["shape"] == [256, 256] and c["count"] == 100][0]
    r = subprocess.run([N.DRIVER_PATH, "--nd", "2", "--size", "256", "--count", "100", "--dump",
                        "--warmup"], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    for line in ("*** Diffusion Benchmark ***", "Grid size: 256x256", "Running the kernel...",
                 "Kernel-only performance:"):
        assert line in out, out
    acc = [l for l in out.splitlines() if l.startswith("Accuracy")][0].split(":")[1].strip()
    assert acc == case["accuracy"], out
    lines = (tmp_path / "diffusion_result.hip.out").read_text().split()
    assert len(lines) == 256 * 256
    samples = np.array(case["samples_bits"], np.uint32).view(np.float32)
    for j, v in zip(range(0, len(lines), case["sample_stride"]), samples):
        assert lines[j] == "%.7E" % float(v), j


def row_tiles(T=None, arith=d3.ARITH_EXACT, dtype=d3.F32):
    return [t for t in d3.tiles() if t["row_stream"] and t["arith"] == arith
            and t["dtype"] == dtype and (T is None or t["temporal_depth"] == T)]


@pytest.mark.parametrize("shape", [(300, 77), (1000, 9), (5, 300), (520, 1), (1, 40)])
def test_every_row_tile_exact_matches_oracle(ctx, shape):
    """2D tiles (every depth, chunking) against the oracle's 2D
    operator on noisy fields, counts that are not multiples of the depth."""
    nx, ny = shape
    k = d3.constants2d_f32(nx, ny)
    f0 = d3.add_noise_f32(d3.initialize2d_f32(nx, ny, k).reshape(-1), 3, 1e-3).reshape(ny, nx)
    want = {c: oracle.run2d_f32(f0, k[8:15], c) for c in (5, 11)}
    c = d3.coefficients(k)
    c[4] = c[5] = 0.0
    for t in row_tiles():
        T = t["temporal_depth"]
        for count in (5, 11):
            for zc in (0, 1, 3):
                print("2D tile", t["index"], "T", T, shape, count, zc, flush=True)
                with d3.Field(ctx, nx, ny, 1) as f:
                    f.set_y_reversed(True)
                    f.upload(f0)
                    tm = f.run(count, c, depth=T, tile=t["index"], zchunks=zc)
                    out = f.download().reshape(ny, nx)
                assert tm["steps"] == count
                assert np.array_equal(out.view(np.uint32), want[count].view(np.uint32)), (t, count, zc)


def test_row_tiles_fma_within_tolerance(ctx):
    nx, ny = 700, 300
    k = d3.constants2d_f32(nx, ny)
    f0 = d3.initialize2d_f32(nx, ny, k)
    want = oracle.run2d_f32(f0, k[8:15], 40)
    c = d3.coefficients(k)
    c[4] = c[5] = 0.0
    for t in row_tiles(arith=d3.ARITH_FMA):
        out = run2d(ctx, f0, c, 40, depth=t["temporal_depth"], tile=t["index"], arith=d3.ARITH_FMA)
        rel, _, _ = oracle.diff_norms(out, want)
        assert rel < 1e-6, (t, rel)


def test_auto_choice_is_row_streaming_and_exact(ctx):
    nx, ny = 2048, 1024
    k = d3.constants2d_f32(nx, ny)
    f0 = d3.initialize2d_f32(nx, ny, k)
    c = d3.coefficients(k)
    for depth in (1, 2, 4, 8):
        with d3.Field(ctx, nx, ny, 1) as f:
            f.set_y_reversed(True)
            f.upload(f0)
            tm = f.run(13, c, depth=depth)
            out = f.download().reshape(ny, nx)
        assert d3.tiles()[tm["tile"]]["row_stream"] == 1, tm
        assert np.array_equal(out.view(np.uint32), oracle.run2d_f32(f0, k[8:15], 13).view(np.uint32))


def test_tuned_2d_entry_runs_the_2d_kernel(tmp_path):
    """Tune-table entries record whether they name a 2D entry: the 2D entry
    (vec 4, 2 groups, 4 rows, 1 wave, prefetch 2) has exactly the parameters
    of a 3D tile, and a table naming it must run stencil2d_kernel, while the
    same parameters flagged 3D run that 3D tile.  An autotuned nz == 1 field
    runs the tile the tuner reported, through a save/load round trip."""
    nx, ny = 2048, 1024
    k = d3.constants2d_f32(nx, ny)
    f0 = d3.initialize2d_f32(nx, ny, k)
    c = d3.coefficients(k)
    want = oracle.run2d_f32(f0, k[8:15], 9)

    def twin(rows):
        return [t for t in d3.tiles() if t["dtype"] == d3.F32 and t["temporal_depth"] == 1
                and t["arith"] == d3.ARITH_EXACT and t["tile_x"] == 512 and t["groups"] == 2
                and t["rows_per_wave"] == 4 and t["waves"] == 1 and t["prefetch"] == 2
                and t["row_stream"] == rows]

    t2d, t3d = twin(1), twin(0)
    assert len(t2d) == 1 and len(t3d) == 1
    for rows, tile in ((1, t2d[0]), (0, t3d[0])):
        path = tmp_path / ("tune%d.txt" % rows)
        # dtype depth fma log2nx log2ny log2nz vec G RPW NW pf zchunks ms depth batch rows
        path.write_text("0 1 0 11 10 0 4 2 4 1 2 0 0.1 1 0 %d\n" % rows)
        with d3.Context(0) as c2, d3.Field(c2, nx, ny, 1) as f:
            c2.load_tuning(str(path))
            f.set_y_reversed(True)
            f.upload(f0)
            tm = f.run(9, c, depth=1)
            out = f.download().reshape(ny, nx)
        assert tm["tile"] == tile["index"], (rows, tm)
        assert np.array_equal(out.view(np.uint32), want.view(np.uint32)), rows
    with d3.Context(0) as c3:
        tuned = c3.autotune(nx, ny, 1, d3.F32, 1)
        assert d3.tiles()[tuned["tile"]]["row_stream"] == 1, tuned
        c3.save_tuning(str(tmp_path / "auto.txt"))
    with d3.Context(0) as c4, d3.Field(c4, nx, ny, 1) as f:
        c4.load_tuning(str(tmp_path / "auto.txt"))
        f.set_y_reversed(True)
        f.upload(f0)
        tm = f.run(9, c, depth=1)
        out = f.download().reshape(ny, nx)
    assert tm["tile"] == tuned["tile"], (tm, tuned)
    assert np.array_equal(out.view(np.uint32), want.view(np.uint32))


def test_row_tiles_refuse_3d_operator(ctx):
    """A 2D tile is only for nz == 1 with ct == cb == 0; nz == 1 fields with
    t/b coefficients run on the 3D tiles and match the 7-point oracle."""
    t = row_tiles(1)[0]
    with d3.Field(ctx, 64, 16, 4) as f:
        f.upload(np.ones((4, 16, 64), np.float32))
        with pytest.raises(d3.D3HError):
            f.run(1, [0.1] * 6 + [0.4], tile=t["index"])
    rng = np.random.default_rng(2)
    f0 = rng.random((1, 40, 300), dtype=np.float32)
    coef = np.array([0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4], np.float32)
    with d3.Field(ctx, 300, 40, 1) as f:
        f.upload(f0)
        with pytest.raises(d3.D3HError):
            f.run(1, coef, tile=t["index"])
        tm = f.run(6, coef)
        assert d3.tiles()[tm["tile"]]["row_stream"] == 0
        out = f.download()
    assert np.array_equal(out.view(np.uint32), oracle.run_f32(f0, coef, 6).view(np.uint32))


def test_full_size_16384_depths_equal_plain(ctx):
    """The 2D bench grid (16384^2): every 2D-kernel depth equals plain
    steps bitwise at full size (device-side comparison)."""
    n = 16384
    k = d3.constants2d_f32(n, n)
    f0 = d3.initialize2d_f32(n, n, k)
    c = d3.coefficients(k)
    with d3.Field(ctx, n, n, 1) as plain:
        plain.set_y_reversed(True)
        plain.upload(f0)
        plain.run(11, c, depth=1)
        for depth in (2, 4, 8):
            with d3.Field(ctx, n, n, 1) as tb:
                tb.set_y_reversed(True)
                tb.upload(f0)
                tm = tb.run(11, c, depth=depth)
                assert d3.tiles()[tm["tile"]]["row_stream"] == 1
                assert tb.compare(plain)["max_abs"] == 0.0, depth


def test_f64_2d_tiles_equal_the_7_point_tiles(ctx):
    """FP64 2D tiles (every depth, chunking) on a row-reversed nz == 1 field
    equal the FP64 7-point tiles with ct = cb = 0 bitwise (positive field, so
    the dropped 0 * c terms cannot change a sign); the 7-point FP64 kernel is
    pinned to the oracle elsewhere."""
    nx, ny = 300, 77
    k = d3.constants2d_f32(nx, ny)
    f0 = d3.add_noise_f32(d3.initialize2d_f32(nx, ny, k).reshape(-1), 5, 1e-3).reshape(ny, nx)
    f0 = f0.astype(np.float64)
    c = d3.coefficients(k)
    c[4] = c[5] = 0.0
    t3 = [t for t in d3.tiles() if t["dtype"] == d3.F64 and t["temporal_depth"] == 1
          and t["arith"] == d3.ARITH_EXACT and not t["row_stream"]][0]
    with d3.Field(ctx, nx, ny, 1, d3.F64) as f:
        f.set_y_reversed(True)
        f.upload(f0)
        f.run(11, c, tile=t3["index"])
        want = f.download()
    tiles = row_tiles(dtype=d3.F64)
    assert tiles
    for t in tiles:
        for zc in (0, 1, 3):
            with d3.Field(ctx, nx, ny, 1, d3.F64) as f:
                f.set_y_reversed(True)
                f.upload(f0)
                tm = f.run(11, c, depth=t["temporal_depth"], tile=t["index"], zchunks=zc)
                got = f.download()
            assert tm["steps"] == 11
            assert np.array_equal(got, want), (t, zc)
