"""CPU-only checks of the C ABI library: it loads, exports every declared
symbol, its host-side field/constants match the reference bitwise, and
argument errors are reported as status codes (no GPU call is made)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import benchmarks_amd as d3
import oracle
from benchmarks_amd import _native as N
from conftest import as_grid, shape_key


def test_library_exports_every_header_symbol():
    names = N.header_functions()
    assert "run_kernel_hip" in names and "d3h_field_run" in names and len(names) >= 40
    handle = ctypes.CDLL(N.LIB_PATH)
    missing = [n for n in names if not hasattr(handle, n)]
    assert not missing, missing
    # and the Python signature table covers the same set
    assert sorted(N.SIGNATURES) == names


def test_ctypes_structs_match_the_header(tmp_path):
    """Every struct the Python binding mirrors has the header's size and field
    offsets (the header compiled by gcc as C)."""
    mirrors = {"d3h_run_config": N.RunConfig, "d3h_timing": N.Timing, "d3h_tile_desc": N.TileDesc,
               "d3h_problem": N.Problem, "d3h_server_config": N.ServerConfig}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "diffusion3d_hip.h"', "int main(void) {"]
    for c, py in mirrors.items():
        lines.append('  printf("%s size %%zu\\n", sizeof(%s));' % (c, c))
        for name, _ in py._fields_:
            lines.append('  printf("%s %s %%zu\\n", offsetof(%s, %s));' % (c, name, c, name))
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-std=c11", "-Wall", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            key, value = line.rsplit(" ", 1)
            got[key] = int(value)
    for c, py in mirrors.items():
        assert got["%s size" % c] == ctypes.sizeof(py), c
        for name, _ in py._fields_:
            assert got["%s %s" % (c, name)] == getattr(py, name).offset, (c, name)


def test_host_constants_match_reference(golden):
    meta, _ = golden
    for key, bits in meta["constants"].items():
        nx, ny, nz = map(int, key.split("x"))
        assert list(d3.constants_f32(nx, ny, nz).view(np.uint32)) == bits, key
        np.testing.assert_array_equal(d3.constants_f64(nx, ny, nz), oracle.constants_f64(nx, ny, nz))


def test_host_initialize_matches_reference(golden):
    meta, arrays = golden
    for shape in [(8, 8, 8), (16, 16, 16), (12, 10, 7), (33, 17, 5), (5, 9, 31)]:
        k = d3.constants_f32(*shape)
        f = d3.initialize_f32(*shape, k)
        ref = as_grid(arrays["out_%s_c0" % shape_key(shape)], shape)
        assert np.array_equal(f.view(np.uint32), ref.view(np.uint32))
    for item in meta["analytical"]:
        shape = tuple(item["shape"])
        k = d3.constants_f32(*shape)
        a = d3.initialize_f32(*shape, k, np.float32(item["count"]) * k[7])
        assert np.array_equal(a.view(np.uint32), as_grid(arrays[item["key"]], shape).view(np.uint32))
    k64 = d3.constants_f64(40, 30, 20)
    np.testing.assert_array_equal(d3.initialize_f64(40, 30, 20, k64, 0.01),
                                  oracle.initialize_f64(40, 30, 20, k64, 0.01))


def test_host_initialize_large_matches_oracle():
    k = d3.constants_f32(128, 128, 128)
    a = d3.initialize_f32(128, 128, 128, k)
    b = oracle.initialize_f32(128, 128, 128, k)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_noise_matches_oracle():
    base = np.linspace(0, 1, 100003, dtype=np.float32)
    a = d3.add_noise_f32(base.copy(), 3, 1e-3)
    b = oracle.add_noise_f32(base, 3, 1e-3)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_tile_table():
    tl = d3.tiles()
    assert len(tl) > 50
    for t in tl:
        assert t["dtype"] in (0, 1, 2, 3) and t["temporal_depth"] in (1, 2, 4, 8)
        assert t["tile_y"] == t["rows_per_wave"] * t["waves"]
        assert t["tile_x"] % 64 == 0
        if t["row_stream"]:               # 2D: streams blocks of rows of nz == 1 fields
            assert t["waves"] == 1 and t["tile_y"] == t["rows_per_wave"] and t["dtype"] in (0, 1)
        else:
            assert t["tile_y"] > 2 * t["temporal_depth"] or t["temporal_depth"] == 1
    for dt in (0, 1, 2, 3):
        for T in (1, 2, 4):
            for ar in (0, 1):
                assert any(t["dtype"] == dt and t["temporal_depth"] == T and t["arith"] == ar
                           and not t["row_stream"] for t in tl), (dt, T, ar)
    for T in (1, 2, 4, 8):
        assert any(t["row_stream"] and t["temporal_depth"] == T for t in tl), T
    with pytest.raises(d3.D3HError):
        N.check(N.lib().d3h_tile_info(len(tl), ctypes.byref(N.TileDesc())), "tile_info")


def test_argument_errors_are_status_codes():
    L = N.lib()
    h = ctypes.c_void_p()
    assert L.d3h_field_create(None, 8, 8, 8, 0, ctypes.byref(h)) == N.D3H_ERR_INVALID
    assert b"NULL" in L.d3h_last_error()
    assert L.d3h_field_run(None, 1, None, None, None) == N.D3H_ERR_INVALID
    assert L.d3h_field_upload(None, None, 0) == N.D3H_ERR_INVALID
    assert L.run_kernel_hip(1, None, 8, 8, 8, *([0.1] * 7)) == N.D3H_ERR_STATE
    assert L.initialize_benchmark_hip(8, 8, 8) == N.D3H_ERR_STATE
    assert L.d3h_constants_f32(0, 8, 8, None) == N.D3H_ERR_INVALID
    assert L.d3h_status_string(N.D3H_ERR_UNSTABLE) == b"unstable coefficients"
    assert L.d3h_context_destroy(None) == N.D3H_OK


def test_driver_help():
    assert os.path.exists(N.DRIVER_PATH)
    r = subprocess.run([N.DRIVER_PATH, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert "--count N" in r.stderr and "--size N" in r.stderr and "--dump" in r.stderr


def test_raw_field_files_round_trip(tmp_path):
    """d3h_raw_write / d3h_raw_read (host side, no GPU): bit-exact round trip
    for every dtype, header fields, and errors for bad or short files."""
    rng = np.random.default_rng(4)
    arrays = [rng.random((3, 5, 7), dtype=np.float32), rng.random((2, 4, 9)),
              rng.integers(0, 65535, (4, 3, 6)).astype(np.uint16),
              rng.random((5, 2, 3)).astype(np.float16)]
    for i, a in enumerate(arrays):
        p = tmp_path / ("f%d.raw" % i)
        d3.save_raw(p, a, steps=17 + i, y_reversed=bool(i % 2))
        b, info = d3.load_raw(p)
        assert b.dtype == a.dtype and b.shape == a.shape
        assert np.array_equal(b.view(np.uint8), a.view(np.uint8))
        assert info["steps"] == 17 + i and info["y_reversed"] == bool(i % 2)
        assert p.stat().st_size == 40 + a.nbytes
    bad = tmp_path / "bad.raw"
    bad.write_bytes(b"NOTAFIELD" + bytes(64))
    with pytest.raises(d3.D3HError):
        d3.load_raw(bad)
    short = tmp_path / "short.raw"
    short.write_bytes((tmp_path / "f0.raw").read_bytes()[:100])
    with pytest.raises(d3.D3HError):
        d3.load_raw(short)
    with pytest.raises(d3.D3HError):
        d3.load_raw(tmp_path / "missing.raw")


def _problem(nx, ny, nz, dtype=N.F32, steps=3, coef=(0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4),
             rounding=0):
    return N.Problem(nx, ny, nz, dtype, steps, (ctypes.c_double * 7)(*coef), rounding, 0)


def test_problem_checks_on_the_host():
    """Failure handling that needs no GPU: unstable or malformed problems come
    back as status codes from the host-side check the batch and server use."""
    L = N.lib()
    assert L.d3h_problem_check(ctypes.byref(_problem(8, 9, 10))) == N.D3H_OK
    bad = [((8, 9, 10), dict(coef=(0.2, 0.2, 0.2, 0.2, 0.2, 0.2, -0.2)), N.D3H_ERR_UNSTABLE),
           ((8, 9, 10), dict(coef=(0.1,) * 6 + (0.5,)), N.D3H_ERR_UNSTABLE),
           ((8, 9, 10), dict(coef=(float("nan"),) + (0.1,) * 5 + (0.5,)), N.D3H_ERR_UNSTABLE),
           ((0, 9, 10), {}, N.D3H_ERR_INVALID),
           ((8, 9, 10), dict(steps=-1), N.D3H_ERR_INVALID),
           ((8, 9, 10), dict(dtype=7), N.D3H_ERR_INVALID),
           ((8, 9, 10), dict(rounding=1), N.D3H_ERR_UNSUPPORTED),       # stochastic + FP32
           ((8, 9, 10), dict(rounding=2, dtype=N.BF16), N.D3H_ERR_INVALID)]
    for dims, kw, code in bad:
        assert L.d3h_problem_check(ctypes.byref(_problem(*dims, **kw))) == code, (dims, kw)
    assert L.d3h_problem_check(ctypes.byref(_problem(8, 9, 10, dtype=N.BF16, rounding=1))) == 0


def test_request_footprint_on_the_host():
    L = N.lib()

    def fp(p, hd, od):
        b = ctypes.c_size_t()
        assert L.d3h_problem_footprint(ctypes.byref(p), hd, od, ctypes.byref(b)) == 0
        return b.value

    def up(v):
        return (v + 255) // 256 * 256
    # FP32, rows already 16-byte multiples: two buffers, no staging
    assert fp(_problem(64, 10, 7), N.F32, N.F32) == 2 * up(64 * 10 * 7 * 4)
    # FP32 with padded rows (nx = 30 -> pitch 32): staging of the dense host size
    assert fp(_problem(30, 10, 7), N.F32, N.F32) == 2 * up(32 * 10 * 7 * 4) + up(30 * 70 * 4)
    # BF16 storage from / to FP32 host arrays: FP32-sized staging
    assert fp(_problem(64, 10, 7, dtype=N.BF16), N.F32, N.F32) == 2 * up(64 * 70 * 2) + up(64 * 70 * 4)
    b = ctypes.c_size_t()
    assert L.d3h_problem_footprint(ctypes.byref(_problem(64, 10, 7, coef=(1,) * 7)), N.F32, N.F32,
                                   ctypes.byref(b)) == N.D3H_ERR_UNSTABLE
