"""GPU parity of stochastic rounding of BF16 storage (D3H_ROUND_STOCHASTIC).

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


def field_run(ctx, bits, coef, counts, seed, **kw):
    nz, ny, nx = bits.shape
    with d3.Field(ctx, nx, ny, nz, d3.BF16) as f:
        f.upload(bits, d3.BF16)
        tms = [f.run(c, coef, rounding=SR, seed=seed, **kw) for c in counts]
        return f.download(), tms


def setup(shape, noise_seed=5):
    k = d3.constants_f32(*shape)
    f0 = d3.add_noise_f32(d3.initialize_f32(*shape, k), noise_seed, 1e-3)
    return oracle.to_lp(f0, 0), np.float32(d3.coefficients(k))


@pytest.mark.parametrize("shape", [(200, 45, 17), (37, 90, 33)])
def test_every_bf16_tile_stochastic_bitwise(ctx, shape):
    bits, coef = setup(shape)
    want = oracle.run_bf16_sr(bits, coef, 11, 123)
    for t in d3.tiles():
        if t["dtype"] != d3.BF16 or t["arith"] != d3.ARITH_EXACT:
            continue
        print("tile", t["index"], "T", t["temporal_depth"], shape, flush=True)
        out, _ = field_run(ctx, bits, coef, [11], 123, depth=t["temporal_depth"], tile=t["index"])
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
    with pytest.raises(d3.D3HError):      # stochastic rounding is BF16-only
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
        k = d3.constants_f32(nx, ny, nz)
        f0 = d3.add_noise_f32(d3.initialize_f32(nx, ny, nz, k), 100 + i, 1e-3)
        bits = oracle.to_lp(f0, 0)
        p = {"nx": nx, "ny": ny, "nz": nz, "dtype": d3.BF16, "steps": steps, "coef": coef,
             "rounding": SR if sr else d3.ROUND_NEAREST, "seed": 1000 + i}
        if sr:
            want = oracle.run_bf16_sr(bits, np.float32(coef), steps, 1000 + i)
        else:
            want = oracle.run_lp(bits, np.float32(coef), steps, 0)
        probs.append(p)
        inits.append(bits)
        wants.append(want)
    for depth in (1, 2, 4):
        with d3.Batch(ctx, probs) as b:
            for i, bits in enumerate(inits):
                b.upload(i, bits, d3.BF16)
            b.run(depth=depth, graph=True)
            for i in range(len(probs)):
                assert np.array_equal(b.download(i, d3.BF16), wants[i]), (depth, i)
    outs = [np.empty_like(w) for w in wants]
    with d3.Server(ctx, hbm_budget_bytes=1 << 28, depth=2, sweep_steps=5) as srv:
        ids = [srv.submit(p, inits[i], d3.BF16, outs[i]) for i, p in enumerate(probs)]
        for r in ids:
            srv.wait(r)
    for i in range(len(probs)):
        assert np.array_equal(outs[i], wants[i]), i
