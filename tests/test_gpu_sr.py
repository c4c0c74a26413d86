"""GPU parity of stochastic rounding of 16-bit storage (D3H_ROUND_STOCHASTIC).

The rounding is a deterministic hash of (seed, cell, step), so every kernel
path -- plain tiles, fused depths, split runs, graph replays, batches and the
request server -- must reproduce the oracle's BF16 stochastic run bitwise.
"""
import numpy as np
import pytest

import benchmarks_amd as d3
import oracle

pytestmark = pytest.mark.gpu

SR = d3.ROUND_STOCHASTIC


def field_run(ctx, bits, coef, counts, seed, dtype=d3.BF16, **kw):
    nz, ny, nx = bits.shape
    with d3.Field(ctx, nx, ny, nz, dtype) as f:
        f.upload(bits if dtype == d3.BF16 else bits.view(np.float16), dtype)
        tms = [f.run(c, coef, rounding=SR, seed=seed, **kw) for c in counts]
        return f.download().view(np.uint16), tms


def setup(shape, noise_seed=5, kind=0):
    k = d3.constants_f32(*shape)
    f0 = d3.add_noise_f32(d3.initialize_f32(*shape, k), noise_seed, 1e-3)
    return oracle.to_lp(f0, kind), np.float32(d3.coefficients(k))


@pytest.mark.parametrize("kind,dtype", [(0, d3.BF16), (1, d3.F16)])
@pytest.mark.parametrize("shape", [(200, 45, 17), (37, 90, 33)])
def test_every_16bit_tile_stochastic_bitwise(ctx, shape, kind, dtype):
    bits, coef = setup(shape, kind=kind)
    want = oracle.run_lp_sr(bits, coef, 11, kind, 123)
    for t in d3.tiles():
        if t["dtype"] != dtype or t["arith"] != d3.ARITH_EXACT:
            continue
        print("tile", t["index"], "T", t["temporal_depth"], shape, flush=True)
        out, _ = field_run(ctx, bits, coef, [11], 123, dtype=dtype, depth=t["temporal_depth"],
                           tile=t["index"])
        assert np.array_equal(out, want), t


def test_split_runs_and_graph_replays_continue_step_numbers(ctx):
    shape = (130, 70, 41)
    bits, coef = setup(shape, 8)
    want = oracle.run_bf16_sr(bits, coef, 13, 7)
    out, tms = field_run(ctx, bits, coef, [6, 7], 7, depth=2)
    assert np.array_equal(out, want)
    # the same 4-step graph replayed from three step offsets
    out, _ = field_run(ctx, bits, coef, [4, 4, 4, 1], 7, depth=4, graph=True)
    assert np.array_equal(out, want)
    # a re-upload restarts the step numbering; another seed changes the result
    out2, _ = field_run(ctx, bits, coef, [13], 8, depth=1)
    assert not np.array_equal(out2, want)
    with pytest.raises(d3.D3HError):      # stochastic rounding is for 16-bit storage
        with d3.Field(ctx, 16, 16, 16, d3.F32) as f:
            f.upload(np.zeros((16, 16, 16), np.float32))
            f.run(1, coef, rounding=SR)


def test_batch_and_server_stochastic_equal_alone(ctx):
    rng = np.random.default_rng(21)
    probs, inits, wants = [], [], []
    for i in range(10):
        nx, ny, nz = [int(v) for v in rng.integers(24, 90, 3)]
        coef, _ = d3.stable_coefficients(nx, ny, nz)
        steps = int(rng.integers(1, 40))
        sr = i % 2 == 0
        kind = 1 if i % 3 == 2 else 0       # FP16 storage for every third problem
        k = d3.constants_f32(nx, ny, nz)
        f0 = d3.add_noise_f32(d3.initialize_f32(nx, ny, nz, k), 100 + i, 1e-3)
        bits = oracle.to_lp(f0, kind)
        p = {"nx": nx, "ny": ny, "nz": nz, "dtype": d3.F16 if kind else d3.BF16, "steps": steps,
             "coef": coef, "rounding": SR if sr else d3.ROUND_NEAREST, "seed": 1000 + i}
        if sr:
            want = oracle.run_lp_sr(bits, np.float32(coef), steps, kind, 1000 + i)
        else:
            want = oracle.run_lp(bits, np.float32(coef), steps, kind)
        probs.append(p)
        inits.append(bits)
        wants.append(want)
    def host(i):        # the host array in the problem's own storage type
        return inits[i].view(np.float16) if probs[i]["dtype"] == d3.F16 else inits[i]

    for depth in (0, 1, 2, 4):
        with d3.Batch(ctx, probs) as b:
            for i in range(len(probs)):
                b.upload(i, host(i), probs[i]["dtype"])
            b.run(depth=depth, graph=True)
            for i in range(len(probs)):
                got = b.download(i, probs[i]["dtype"]).view(np.uint16)
                assert np.array_equal(got, wants[i]), (depth, i)
    outs = [np.empty_like(host(i)) for i in range(len(probs))]
    with d3.Server(ctx, hbm_budget_bytes=1 << 28, depth=2, sweep_steps=5) as srv:
        ids = [srv.submit(p, host(i), p["dtype"], outs[i]) for i, p in enumerate(probs)]
        for r in ids:
            srv.wait(r)
    for i in range(len(probs)):
        assert np.array_equal(outs[i].view(np.uint16), wants[i]), i
