"""Pins the CPU oracle (oracle/) before anything is checked against it.

* Bitwise against the reference's own CPU build (fixtures in tests/golden,
  written by tests/golden/gen_golden.py from oracle/_ref/libref.so).
* Closed forms and invariants the mathematics fixes (coefficients for cubic
  grids, conservation, maximum principle, O(dx^2) convergence to the exact
  PDE solution), and an independent numpy restatement on tiny grids.
* Planted mistakes (swapped y terms, a dropped term, a contracted FMA) are
  shown to change the fixture bits, so the fixtures do pin the term order.
"""
import hashlib
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, as_grid, shape_key

FULL_SHAPES = [(8, 8, 8), (16, 16, 16), (12, 10, 7), (33, 17, 5), (5, 9, 31)]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_constants_bitwise(golden):
    meta, _ = golden
    for key, bits in meta["constants"].items():
        nx, ny, nz = map(int, key.split("x"))
        got = oracle.constants_f32(nx, ny, nz).view(np.uint32)
        assert list(got) == bits, key


@pytest.mark.parametrize("n", [8, 16, 32, 128, 256, 512, 768])
def test_constants_closed_form(n):
    k = oracle.constants_f32(n, n, n)
    dx = np.float32(1.0) / np.float32(n)
    assert k[1] == dx
    assert k[7] == dx * dx                      # dt = 0.1 dx^2 / kappa rounds to dx^2
    assert k[8] == np.float32(0.1) and k[9] == k[8] and k[10] == k[8] and k[12] == k[8]
    assert k[14] == np.float32(0.39999998)      # cc = 1 - 6 ce in float/double as the reference
    assert k[4] == np.float32(2 * np.pi)


def test_initial_and_analytical_fields_bitwise(golden):
    meta, arrays = golden
    for shape in FULL_SHAPES:
        k = oracle.constants_f32(*shape)
        f0 = oracle.initialize_f32(*shape, k, 0.0)
        ref = as_grid(arrays["out_%s_c0" % shape_key(shape)], shape)
        assert np.array_equal(f0.view(np.uint32), ref.view(np.uint32)), shape
    for item in meta["analytical"]:
        shape = tuple(item["shape"])
        k = oracle.constants_f32(*shape)
        a = oracle.initialize_f32(*shape, k, np.float32(item["count"]) * k[7])
        ref = as_grid(arrays[item["key"]], shape)
        assert np.array_equal(a.view(np.uint32), ref.view(np.uint32)), item["key"]


def test_run_bitwise_full_arrays(golden):
    meta, arrays = golden
    for case in meta["cases"]:
        shape = tuple(case["shape"])
        if shape not in FULL_SHAPES:
            continue
        k = oracle.constants_f32(*shape)
        f = oracle.run_f32(oracle.initialize_f32(*shape, k), k[8:15], case["count"])
        ref = as_grid(arrays["out_%s_c%d" % (shape_key(shape), case["count"])], shape)
        assert np.array_equal(f.view(np.uint32), ref.view(np.uint32)), case
        acc = oracle.accuracy_f32(f, k, case["count"])
        assert np.float32(acc).view(np.uint32) == case["accuracy_bits"], case


def test_run_bitwise_hashed(golden):
    meta, _ = golden
    for case in meta["cases"]:
        shape = tuple(case["shape"])
        if shape in FULL_SHAPES:
            continue
        k = oracle.constants_f32(*shape)
        f = oracle.run_f32(oracle.initialize_f32(*shape, k), k[8:15], case["count"])
        assert sha(f) == case["sha256"], case
        flat = f.reshape(-1)
        idx = np.arange(0, flat.size, case["sample_stride"])
        assert list(flat[idx].view(np.uint32)) == case["samples_bits"]
        assert np.float32(oracle.accuracy_f32(f, k, case["count"])).view(np.uint32) == \
            case["accuracy_bits"]


def test_reference_default_known_answer(golden):
    """256^3, 100 steps (the reference's defaults, diffusion3d/main.cc:65-67,131)."""
    meta, _ = golden
    k = oracle.constants_f32(256, 256, 256)
    f = oracle.run_f32(oracle.initialize_f32(256, 256, 256, k), k[8:15], 100)
    assert sha(f) == meta["default_256_c100"]["sha256"]
    acc = oracle.accuracy_f32(f, k, 100)
    assert "%e" % acc == meta["default_256_c100"]["accuracy"] == "4.336409e-07"


def test_openmp_temporal_blocking_fixture_equals_plain(golden):
    """The reference's OpenMP temporal blocking is bitwise equal to its plain path
    (diffusion3d/diffusion3d_openmp_temporal_blocking.cc), for 2, 4 and 8 threads."""
    meta, _ = golden
    plain = {c["count"]: c["sha256"] for c in meta["cases"] if tuple(c["shape"]) == (32, 32, 32)}
    for tb in meta["tb"]:
        assert tb["sha256"] == plain[tb["count"]], tb


def _numpy_step(f, c, order="reference", fma=False, dtype=np.float32):
    """Independent restatement with edge-replicated padding (== index clamping)."""
    f = f.astype(dtype)
    p = np.pad(f, 1, mode="edge")
    C = p[1:-1, 1:-1, 1:-1]
    W, E = p[1:-1, 1:-1, :-2], p[1:-1, 1:-1, 2:]
    N, S = p[1:-1, :-2, 1:-1], p[1:-1, 2:, 1:-1]
    B, T = p[:-2, 1:-1, 1:-1], p[2:, 1:-1, 1:-1]
    ce, cw, cn, cs, ct, cb, cc = [dtype(v) for v in c]
    if fma:
        # fma(a,b,c) in float32 via float64 (exact product, one rounding)
        def fm(a, b, z):
            return (a.astype(np.float64) * b.astype(np.float64) + z.astype(np.float64)).astype(dtype)
        r = fm(np.full_like(C, cc), C, cw * W)
        for kk, v in ((ce, E), (cs, S), (cn, N), (cb, B), (ct, T)):
            r = fm(np.full_like(C, kk), v, r)
        return r
    if order == "reference":
        return cc * C + cw * W + ce * E + cs * S + cn * N + cb * B + ct * T
    if order == "swap_y":
        return cc * C + cw * W + ce * E + cn * N + cs * S + cb * B + ct * T
    if order == "drop_t":
        return cc * C + cw * W + ce * E + cs * S + cn * N + cb * B
    raise ValueError(order)


def test_numpy_restatement_matches_oracle_and_fixture(golden):
    _, arrays = golden
    shape = (12, 10, 7)
    k = oracle.constants_f32(*shape)
    f = oracle.initialize_f32(*shape, k)
    for _ in range(7):
        f = _numpy_step(f, k[8:15])
    ref = as_grid(arrays["out_12x10x7_c7"], shape)
    assert np.array_equal(f.view(np.uint32), ref.view(np.uint32))
    # FP64: numpy restatement == oracle_run_f64, bitwise
    k64 = oracle.constants_f64(*shape)
    g = oracle.initialize_f64(*shape, k64)
    h = g.copy()
    for _ in range(5):
        h = _numpy_step(h, k64[8:15], dtype=np.float64)
    assert np.array_equal(h, oracle.run_f64(g, k64[8:15], 5))


@pytest.mark.parametrize("mistake", ["swap_y", "drop_t", "fma"])
def test_planted_mistakes_change_the_bits(golden, mistake):
    _, arrays = golden
    shape = (33, 17, 5)
    k = oracle.constants_f32(*shape)
    f = oracle.initialize_f32(*shape, k)
    for _ in range(7):
        f = _numpy_step(f, k[8:15], order=mistake if mistake != "fma" else "reference",
                        fma=(mistake == "fma"))
    ref = as_grid(arrays["out_33x17x5_c7"], shape)
    assert not np.array_equal(f.view(np.uint32), ref.view(np.uint32))


def test_brute_force_tiny_grid_fp64():
    """Hand-indexed loops with explicit clamping on a 3x4x5 grid."""
    nx, ny, nz = 3, 4, 5
    rng = np.random.default_rng(0)
    f = rng.random((nz, ny, nx))
    c = [0.11, 0.12, 0.13, 0.14, 0.15, 0.16, 0.19]
    ce, cw, cn, cs, ct, cb, cc = c
    g = np.empty_like(f)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                w = f[z, y, max(x - 1, 0)]
                e = f[z, y, min(x + 1, nx - 1)]
                n = f[z, max(y - 1, 0), x]
                s = f[z, min(y + 1, ny - 1), x]
                b = f[max(z - 1, 0), y, x]
                t = f[min(z + 1, nz - 1), y, x]
                g[z, y, x] = cc * f[z, y, x] + cw * w + ce * e + cs * s + cn * n + cb * b + ct * t
    assert np.array_equal(oracle.run_f64(f, c, 1), g)


def test_conservation_and_maximum_principle_fp64():
    shape = (20, 13, 9)
    k = oracle.constants_f64(*shape)
    rng = np.random.default_rng(1)
    f0 = rng.random(shape[::-1])
    f = oracle.run_f64(f0, k[8:15], 50)
    # clamped boundaries conserve the sum exactly in exact arithmetic
    assert abs(f.sum() - f0.sum()) / f0.sum() < 1e-12
    assert f.min() >= f0.min() - 1e-15 and f.max() <= f0.max() + 1e-15


def test_fp32_mass_drift_matches_coefficient_deficit():
    """cc + 6 ce - 1 = -1.49e-8 in float; the sum shrinks by about that per step."""
    n = 32
    k = oracle.constants_f32(n, n, n)
    f0 = oracle.initialize_f32(n, n, n, k)
    f = oracle.run_f32(f0, k[8:15], 100)
    deficit = float(np.float64(k[14]) + 6 * np.float64(k[8]) - 1.0)
    drift = (f.astype(np.float64).sum() - f0.astype(np.float64).sum()) / f0.astype(np.float64).sum()
    assert abs(drift - 100 * deficit) < 3e-6


def test_second_order_convergence_fp64():
    """RMS error vs the exact solution falls ~4x per halving of dx at fixed t."""
    errs = []
    for n, count in ((16, 20), (32, 80), (64, 320)):
        k = oracle.constants_f64(n, n, n)
        f = oracle.run_f64(oracle.initialize_f64(n, n, n, k), k[8:15], count)
        errs.append(oracle.rms_vs_analytical_f64(f, k, count))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.5 < r1 < 4.5 and 3.7 < r2 < 4.3, errs


def test_f16_conversion_matches_numpy():
    rng = np.random.default_rng(2)
    x = np.concatenate([
        rng.standard_normal(20000).astype(np.float32),
        (rng.random(20000) * 1e-4).astype(np.float32),          # subnormal halves
        (rng.random(2000) * 1e-7).astype(np.float32),
        np.array([0.0, -0.0, 65504.0, 65520.0, 1e6, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26],
                 np.float32),
    ])
    # exact halfway cases between adjacent halves
    h = rng.integers(0, 0x7bff, 5000).astype(np.uint16).view(np.float16).astype(np.float32)
    h2 = (h.view(np.uint32) + (1 << 12)).view(np.float32)
    x = np.concatenate([x, h2])
    got = oracle.to_lp(x, 1)
    with np.errstate(over="ignore"):  # 65520 and 1e6 round to inf on purpose
        want = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)
    back = oracle.from_lp(want, 1)
    assert np.array_equal(back, want.view(np.float16).astype(np.float32))


def test_bf16_conversion_round_to_nearest_even():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(50000).astype(np.float32),
                        (rng.random(5000) * 1e-3).astype(np.float32)])
    # plant exact ties: low 16 bits = 0x8000
    ties = ((rng.integers(0x3c00, 0x4200, 5000).astype(np.uint32) << 16) | 0x8000).view(np.float32)
    x = np.concatenate([x, ties])
    got = oracle.to_lp(x, 0).astype(np.uint32)
    u = x.view(np.uint32)
    lo = u & 0xffff0000
    hi = lo + 0x10000
    vlo = lo.view(np.float32).astype(np.float64)
    vhi = hi.view(np.float32).astype(np.float64)
    xv = x.astype(np.float64)
    dlo, dhi = np.abs(xv - vlo), np.abs(vhi - xv)
    pick_hi = (dhi < dlo) | ((dhi == dlo) & (((lo >> 16) & 1) == 1))
    want = np.where(pick_hi, hi, lo) >> 16
    assert np.array_equal(got, want)


def test_noise_properties():
    n = 1 << 18
    a = oracle.add_noise_f32(np.zeros(n, np.float32), 3, 1e-3)
    b = oracle.add_noise_f32(np.zeros(n, np.float32), 3, 1e-3)
    c = oracle.add_noise_f32(np.zeros(n, np.float32), 4, 1e-3)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= -1e-3 and a.max() < 1e-3
    assert abs(a.mean()) < 1e-5 and abs(a.std() - 1e-3 / np.sqrt(3)) < 1e-5


# ---- stochastic rounding of BF16 storage (no reference counterpart) --------

def _lowbias32_np(x):
    x = np.asarray(x, dtype=np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7feb352d) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846ca68b) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def _sr_hash_np(seed, row, pair, step):
    """Independent restatement: lowbias32 of the row and step, plus a Weyl
    increment per x pair."""
    row = np.asarray(row, dtype=np.uint64)
    pair = np.asarray(pair, dtype=np.uint64)
    base = _lowbias32_np((row * 0x9E3779B1 + step * 0x85EBCA77 + seed) & 0xFFFFFFFF)
    return (base + pair * 0x9E3779B9) & 0xFFFFFFFF


def _sr_bits_np(shape2d, step, seed):
    """The 16 random bits of every cell of an (rows, nx) array."""
    rows, nx = shape2d
    x = np.arange(nx, dtype=np.uint64)
    h = _sr_hash_np(seed, np.arange(rows, dtype=np.uint64)[:, None], x[None, :] // 2, step)
    return np.where(x[None, :] & 1, h & 0xFFFF, h >> 16)


def test_sr_hash_matches_independent_restatement():
    rng = np.random.default_rng(0)
    for seed, row, pair, step in zip(rng.integers(0, 2**32, 50), rng.integers(0, 2**32, 50),
                                     rng.integers(0, 2**31, 50), rng.integers(0, 2**31, 50)):
        assert oracle.sr_hash(int(seed), int(row), int(pair), int(step)) == int(
            _sr_hash_np(int(seed), int(row), int(pair), int(step)))
    # along a row, both halves take every value about equally often and
    # neighbouring cells' bits are unrelated
    h = np.array([oracle.sr_hash(7, 11, c, 3) for c in range(4096)], np.uint64)
    for half in (h & 0xFFFF, h >> 16):
        assert 0.45 < (half < 32768).mean() < 0.55 and len(set(half.tolist())) > 3900
    assert abs(np.corrcoef((h & 0xFFFF).astype(float), (h >> 16).astype(float))[0, 1]) < 0.05
    # one cell over consecutive steps, and one column over consecutive rows
    for seq in (np.array([oracle.sr_hash(7, 11, 5, t) for t in range(4096)], np.uint64),
                np.array([oracle.sr_hash(7, y, 5, 3) for y in range(4096)], np.uint64)):
        for half in (seq & 0xFFFF, seq >> 16):
            v = half.astype(float)
            assert 0.45 < (v < 32768).mean() < 0.55
            assert abs(np.corrcoef(v[:-1], v[1:])[0, 1]) < 0.05


def _sr_bf16_np(f, step, seed):
    """BF16 stochastic rounding of an (nz, ny, nx) float32 array (finite
    values): the cell's 16 random bits added below the mantissa, truncated."""
    nx = f.shape[-1]
    u = f.view(np.uint32).reshape(-1, nx).astype(np.uint64)
    r = _sr_bits_np(u.shape, step, seed)
    return (((u + r) >> 16) & 0xFFFF).astype(np.uint16).reshape(f.shape)


def test_sr_round_matches_gfx950_conversions():
    """oracle_sr_round against outputs of gfx950's v_cvt_sr_bf16_f32 and
    v_cvt_sr_f16_f32 recorded on an MI355X (tests/golden/gen_sr_hw.py): the
    instruction takes the high 16 bits of its random operand, so cell bits r
    are the operand r << 16.  Zeros, subnormals, overflow edges, inf and nan
    payloads included."""
    v = np.load(os.path.join(GOLDEN, "sr_hw_vectors.npz"))
    x, r = v["x"], v["r"]
    assert len(x) > 18000
    fx = x.view(np.float32)
    for kind, want in ((0, v["bf16"]), (1, v["f16"])):
        got = np.array([oracle.sr_round(fx[i], int(r[i] >> 16), kind) for i in range(len(x))],
                       np.uint16)
        bad = np.nonzero(got != want)[0]
        assert len(bad) == 0, [(hex(x[i]), hex(r[i]), hex(got[i]), hex(want[i])) for i in bad[:5]]


def test_bf16_sr_run_matches_numpy_restatement():
    shape = (9, 7, 5)                      # (nz, ny, nx) of a 5x7x9 grid
    rng = np.random.default_rng(3)
    f0 = rng.random(shape, dtype=np.float32)
    bits = oracle.to_lp(f0, 0)
    c = np.array([0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4], np.float32)
    want = bits.copy()
    for t in range(6):
        a = oracle.from_lp(want, 0)
        want = _sr_bf16_np(oracle.run_f32(a, c, 1), 10 + t, 99)
    got = oracle.run_bf16_sr(bits, c, 6, 99, step0=10)
    assert np.array_equal(got, want)
    # split runs continue the step numbering; an off-by-one step changes bits
    half = oracle.run_bf16_sr(oracle.run_bf16_sr(bits, c, 2, 99, step0=10), c, 4, 99, step0=12)
    assert np.array_equal(half, want)
    assert not np.array_equal(oracle.run_bf16_sr(bits, c, 6, 99, step0=11), want)


def test_bf16_sr_rounds_to_a_neighbour_without_bias():
    rng = np.random.default_rng(4)
    x = ((rng.random(200000) * 2 - 1).astype(np.float32) * np.float32(0.3)).reshape(1, 1, -1)
    r = oracle.from_lp(_sr_bf16_np(x, 0, 5), 0)
    lo = (x.view(np.uint32) & 0xFFFF0000).view(np.float32)        # toward zero
    hi = ((x.view(np.uint32) & 0xFFFF0000) + 0x10000).view(np.float32)
    assert np.all((r == lo) | (r == hi))
    # exactly representable values are kept
    xr = oracle.from_lp(oracle.to_lp(x, 0), 0)
    assert np.array_equal(oracle.from_lp(_sr_bf16_np(xr, 1, 5), 0), xr)
    # unbiased: one value rounded at many cells (one row and many rows)
    # averages back to itself
    ulp = 2.0 ** -11                                                 # bf16 ulp at 0.12
    for shape in ((1, 1, 400000), (1, 400000, 1)):
        v = np.full(shape, np.float32(0.1234567), np.float32)
        m = oracle.from_lp(_sr_bf16_np(v, 2, 9), 0).astype(np.float64).mean()
        assert abs(m - 0.1234567) < 5 * ulp / np.sqrt(v.size), shape
    # and one cell rounded at many steps averages back to itself
    r = np.array([oracle.sr_hash(9, 4, 3, t) >> 16 for t in range(100000)], np.uint64)
    u = np.uint64(np.float32(0.1234567).view(np.uint32))
    m = ((((u + r) >> 16) << 16).astype(np.uint32).view(np.float32)).astype(np.float64).mean()
    assert abs(m - 0.1234567) < 5 * ulp / np.sqrt(r.size)


def test_bf16_sr_tracks_fp64_where_nearest_stalls():
    shape = (48, 40, 56)
    nx, ny, nz = shape
    h = 1.0 / max(shape)
    dt = 0.1 * h * h / 0.1
    ce, cn, ct = (np.float32(0.1 * dt * n * n) for n in shape)
    c = np.array([ce, ce, cn, cn, ct, ct, 1 - 2 * (ce + cn + ct)], np.float32)
    k = oracle.constants_f32(*shape)
    f0 = oracle.add_noise_f32(oracle.initialize_f32(*shape, k), 7, 1e-3)
    ref = oracle.run_f64(f0.astype(np.float64), c.astype(np.float64), 200)
    bits = oracle.to_lp(f0, 0)
    rne = oracle.from_lp(oracle.run_lp(bits, c, 200, 0), 0)
    sr = oracle.from_lp(oracle.run_bf16_sr(bits, c, 200, 1), 0)
    def rel(a):
        return np.linalg.norm(a - ref) / np.linalg.norm(ref)
    print("bf16 rel L2 vs fp64: nearest %.3e stochastic %.3e" % (rel(rne), rel(sr)))
    assert rel(rne) > 0.05 and rel(sr) < 5e-3


@pytest.mark.parametrize("shape", [(16, 16, 16), (12, 10, 7), (64, 32, 16)])
def test_fp64_constants_and_field_closed_form(shape):
    """The FP64 instantiation (no reference build): constants from their
    formulas, and the cosine product against numpy's float64 evaluation."""
    nx, ny, nz = shape
    k = oracle.constants_f64(*shape)
    dx, dy, dz = 1.0 / nx, 1.0 / ny, 1.0 / nz
    dt = 0.1 * dx * dx / 0.1
    want = [0.1, dx, dy, dz, 2 * np.pi, 2 * np.pi, 2 * np.pi, dt, 0.1 * dt / dx ** 2,
            0.1 * dt / dx ** 2, 0.1 * dt / dy ** 2, 0.1 * dt / dy ** 2, 0.1 * dt / dz ** 2,
            0.1 * dt / dz ** 2]
    assert np.allclose(k[:14], want, rtol=1e-15, atol=0)
    assert k[14] == 1.0 - (k[8] + k[9] + k[10] + k[11] + k[12] + k[13])
    t = 37 * dt
    f = oracle.initialize_f64(*shape, k, t)
    a = np.exp(-0.1 * t * (2 * np.pi) ** 2)
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    z = (np.arange(nz) + 0.5) * dz
    ref = (0.125 * (1 - a * np.cos(2 * np.pi * x))[None, None, :]
           * (1 - a * np.cos(2 * np.pi * y))[None, :, None]
           * (1 - a * np.cos(2 * np.pi * z))[:, None, None])
    assert np.max(np.abs(f - ref)) < 4e-17


def _sr_f16_np(f, step, seed):
    """FP16 stochastic rounding of an (nz, ny, nx) float32 array (finite
    values) -> float32 values, in float64 arithmetic: |x| + (r >> 3) ulp / 2^13
    truncated to the FP16 grid (ulp 2^-24 below 2^-14), sign kept, >= 2^16 to
    inf."""
    nx = f.shape[-1]
    u = f.view(np.uint32).reshape(-1, nx)
    rr = (_sr_bits_np(u.shape, step, seed) >> 3).astype(np.float64)
    a = np.abs(f.reshape(-1, nx).astype(np.float64))
    e = ((u >> 23) & 0xFF).astype(np.int64)
    ulp = np.where(e < 113, 2.0 ** -24, np.ldexp(1.0, np.maximum(e, 113) - 137))
    q = np.floor((a + rr * ulp / 8192.0) / ulp) * ulp
    q = np.where(q >= 65536.0, np.inf, q)
    return np.copysign(q, f.reshape(-1, nx).astype(np.float64)).astype(np.float32).reshape(f.shape)


def test_f16_sr_run_matches_numpy_restatement():
    rng = np.random.default_rng(8)
    f0 = (rng.random((6, 9, 7), dtype=np.float32) * np.float32(0.5)).astype(np.float32)
    f0[0, 0, :3] = [0.0, 3e-6, -2e-7]            # zero and FP16 subnormals (stochastic too)
    bits = oracle.to_lp(f0, 1)
    c = np.array([0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4], np.float32)
    a = oracle.from_lp(bits, 1)
    with np.errstate(over="ignore"):
        for t in range(5):
            a = _sr_f16_np(oracle.run_f32(a, c, 1), 3 + t, 17)
    got = oracle.from_lp(oracle.run_lp_sr(bits, c, 5, 1, 17, step0=3), 1)
    assert np.array_equal(got.view(np.uint32), a.view(np.uint32))
    # kind 0 through the general entry equals the BF16 function
    b16 = oracle.to_lp(f0, 0)
    assert np.array_equal(oracle.run_lp_sr(b16, c, 4, 0, 5), oracle.run_bf16_sr(b16, c, 4, 5))


def test_f16_sr_tracks_fp64_where_nearest_stalls():
    shape = (48, 40, 56)
    h = 1.0 / max(shape)
    dt = 0.1 * h * h / 0.1
    ce, cn, ct = (np.float32(0.1 * dt * n * n) for n in shape)
    c = np.array([ce, ce, cn, cn, ct, ct, 1 - 2 * (ce + cn + ct)], np.float32)
    k = oracle.constants_f32(*shape)
    f0 = oracle.add_noise_f32(oracle.initialize_f32(*shape, k), 7, 1e-3)
    ref = oracle.run_f64(f0.astype(np.float64), c.astype(np.float64), 400)
    bits = oracle.to_lp(f0, 1)
    rne = oracle.from_lp(oracle.run_lp(bits, c, 400, 1), 1)
    sr = oracle.from_lp(oracle.run_lp_sr(bits, c, 400, 1, 3), 1)
    def rel(a):
        return np.linalg.norm(a - ref) / np.linalg.norm(ref)
    print("f16 rel L2 vs fp64: nearest %.3e stochastic %.3e" % (rel(rne), rel(sr)))
    assert rel(sr) < rel(rne) / 5 and rel(sr) < 1e-3
