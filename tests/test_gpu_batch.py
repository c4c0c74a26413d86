"""GPU parity of batched execution: every member of a batch of uneven grids
(mixed dtypes and step counts) ends bitwise equal to the oracle run of that
problem alone (EXACT mode), with and without graph capture, for each depth."""
import numpy as np
import pytest

import benchmarks_amd as d3
import oracle

pytestmark = pytest.mark.gpu


def make_problems(n, seed, lo=8, hi=140, max_steps=30, dtypes=(d3.F32, d3.BF16, d3.F16, d3.F64)):
    rng = np.random.default_rng(seed)
    probs, inits = [], []
    for i in range(n):
        nx, ny, nz = [int(np.exp(rng.uniform(np.log(lo), np.log(hi)))) for _ in range(3)]
        dt = dtypes[i % len(dtypes)]
        coef, _ = d3.stable_coefficients(nx, ny, nz, dtype=np.float64 if dt == d3.F64 else np.float32)
        steps = int(rng.integers(0, max_steps + 1))
        k = d3.constants_f32(nx, ny, nz)
        f0 = d3.add_noise_f32(d3.initialize_f32(nx, ny, nz, k), 100 + i, 1e-3)
        probs.append({"nx": nx, "ny": ny, "nz": nz, "dtype": dt, "steps": steps, "coef": coef})
        inits.append(f0)
    return probs, inits


def expected(p, f0):
    if p["dtype"] == d3.F32:
        return oracle.run_f32(f0, np.float32(p["coef"]), p["steps"])
    if p["dtype"] == d3.F64:
        return oracle.run_f64(f0.astype(np.float64), p["coef"], p["steps"])
    kind = 0 if p["dtype"] == d3.BF16 else 1
    return oracle.run_lp(oracle.to_lp(f0, kind), np.float32(p["coef"]), p["steps"], kind)


def bits(a):
    return a.view(np.uint16) if a.dtype == np.float16 else a


@pytest.mark.parametrize("depth,graph", [(1, False), (2, True), (4, True), (2, False)])
def test_batch_members_equal_oracle(ctx, depth, graph):
    probs, inits = make_problems(24, seed=depth * 10 + graph)
    want = [expected(p, f) for p, f in zip(probs, inits)]
    with d3.Batch(ctx, probs) as b:
        for i, f in enumerate(inits):
            b.upload(i, f, d3.F32)
        t = b.run(depth=depth, graph=graph)
        assert t["launches"] > 0
        for i, p in enumerate(probs):
            got = b.download(i)
            assert np.array_equal(bits(got), bits(want[i])), (i, p)


def test_batch_graph_replay_and_alone_vs_mix(ctx):
    probs, inits = make_problems(12, seed=7, dtypes=(d3.F32,))
    with d3.Batch(ctx, probs) as b:
        for rep in range(2):
            for i, f in enumerate(inits):
                b.upload(i, f, d3.F32)
            b.run(depth=2, graph=True)
            mix = [b.download(i) for i in range(len(probs))]
    for i, p in enumerate(probs):
        with d3.Field(ctx, p["nx"], p["ny"], p["nz"], d3.F32) as f:
            f.upload(inits[i])
            f.run(p["steps"], p["coef"])
            alone = f.download()
        assert np.array_equal(alone.view(np.uint32), mix[i].view(np.uint32)), i


def test_batch_rejects_unstable_coefficients(ctx):
    k = d3.constants_f32(32, 64, 64)   # reference dt rule (dx only) on a slab: cc < 0
    assert k[14] < 0
    p = {"nx": 32, "ny": 64, "nz": 64, "dtype": d3.F32, "steps": 1, "coef": k[8:15]}
    with pytest.raises(d3.D3HError) as e:
        d3.Batch(ctx, [p])
    assert e.value.status == 7


def test_whole_batch_transfers_equal_per_problem(ctx):
    """upload_all / download_all (one synchronisation, per-problem staging
    slots) move the same bytes as the per-problem calls, for pinned and
    pageable buffers and a converting host dtype."""
    from benchmarks_amd import workloads
    rng = np.random.default_rng(31)
    probs, inits = [], []
    for i in range(9):
        nx, ny, nz = [int(v) for v in rng.integers(5, 70, 3)]
        coef, _ = d3.stable_coefficients(nx, ny, nz)
        dt = [d3.F32, d3.BF16, d3.F64][i % 3]
        probs.append({"nx": nx, "ny": ny, "nz": nz, "dtype": dt, "steps": int(rng.integers(1, 9)),
                      "coef": coef})
        k = d3.constants_f32(nx, ny, nz)
        inits.append(d3.add_noise_f32(d3.initialize_f32(nx, ny, nz, k), 50 + i, 1e-3))
    arena = workloads.PinnedArena()
    try:
        pinned = []
        for f in inits:
            a = arena.empty(f.shape)
            a[...] = f
            pinned.append(a)
        with d3.Batch(ctx, probs) as a, d3.Batch(ctx, probs) as b:
            for i, f in enumerate(inits):
                a.upload(i, f, d3.F32)
            b.upload_all(pinned, d3.F32)
            a.run(depth=2)
            b.run(depth=2)
            outs = [arena.empty(f.shape) for f in inits]
            b.download_all(outs, d3.F32)
            pageable = b.download_all([np.empty_like(f) for f in inits], d3.F32)
            for i in range(len(probs)):
                want = a.download(i, d3.F32)
                assert np.array_equal(outs[i].view(np.uint32), want.view(np.uint32)), i
                assert np.array_equal(pageable[i].view(np.uint32), want.view(np.uint32)), i
            with pytest.raises(ValueError):
                b.download_all(outs[:-1], d3.F32)
    finally:
        arena.close()


def test_batch_autotune_choices_keep_results(tmp_path):
    """The batch autotuner's per-size-class choices (depth, tile, z-chunks,
    all dtypes) change the launches, not the results: depth-0 runs after
    tuning equal the oracle bitwise; the saved table holds batch entries and
    a context that loads it produces the same bits."""
    probs, inits = make_problems(20, seed=77, lo=8, hi=90, max_steps=13)
    want = [expected(p, f) for p, f in zip(probs, inits)]
    path = str(tmp_path / "batch_tune.txt")
    with d3.Context(0) as ctx:
        with d3.Batch(ctx, probs) as b:
            info = b.autotune()
            print("batch autotune:", info, flush=True)
            assert info["classes"] >= 1
            assert info["sweep_ms_after"] > 0
            b.upload_all(inits, d3.F32)
            b.run(depth=0, graph=True)
            for i, p in enumerate(probs):
                assert np.array_equal(bits(b.download(i)), bits(want[i])), (i, p)
        ctx.save_tuning(path)
    # columns: dtype depth fma log2nx log2ny log2nz vec G RPW NW pf zchunks ms depth batch rows
    rows = [l.split() for l in open(path) if not l.startswith("#")]
    assert any(r[14] == "1" for r in rows)
    assert all(r[15] == "0" for r in rows if r[14] == "1")
    with d3.Context(0) as ctx2:
        ctx2.load_tuning(path)
        with d3.Batch(ctx2, probs) as b:
            b.upload_all(inits, d3.F32)
            b.run(depth=0, graph=False)
            for i, p in enumerate(probs):
                assert np.array_equal(bits(b.download(i)), bits(want[i])), (i, p)
