"""GPU tests of the request server: requests of mixed size, step count and
storage type are submitted from several threads while others are running;
each result must equal the oracle run of that request alone, bitwise (EXACT
mode), whatever else shared its sweeps."""
import threading
import time

import numpy as np
import pytest

import benchmarks_amd as d3
import oracle
from benchmarks_amd import _native as N

pytestmark = pytest.mark.gpu


def make_request(rng, i, lo=12, hi=90, max_steps=40):
    nx, ny, nz = [int(np.exp(rng.uniform(np.log(lo), np.log(hi)))) for _ in range(3)]
    dt = d3.BF16 if i % 3 == 2 else d3.F32
    coef, _ = d3.stable_coefficients(nx, ny, nz)
    steps = int(rng.integers(0, max_steps + 1))
    k = d3.constants_f32(nx, ny, nz)
    f0 = d3.add_noise_f32(d3.initialize_f32(nx, ny, nz, k), 1000 + i, 1e-3)
    return {"nx": nx, "ny": ny, "nz": nz, "dtype": dt, "steps": steps, "coef": coef}, f0


def expected(p, f0):
    if p["dtype"] == d3.F32:
        return oracle.run_f32(f0, np.float32(p["coef"]), p["steps"])
    # BF16 storage, result returned as float32
    bits = oracle.run_lp(oracle.to_lp(f0, 0), np.float32(p["coef"]), p["steps"], 0)
    return oracle.from_lp(bits, 0)


@pytest.mark.parametrize("depth,sweep", [(2, 8), (1, 5), (4, 16)])
def test_server_results_equal_alone(ctx, depth, sweep):
    rng = np.random.default_rng(depth * 100 + sweep)
    reqs = [make_request(rng, i) for i in range(30)]
    outs = [np.empty(f0.shape, np.float32) for _, f0 in reqs]
    ids = [None] * len(reqs)
    with d3.Server(ctx, hbm_budget_bytes=1 << 30, depth=depth, sweep_steps=sweep) as srv:
        def client(part):
            for i in part:
                p, f0 = reqs[i]
                ids[i] = srv.submit(p, f0, d3.F32, outs[i])
                time.sleep(0.001 * (i % 3))
        threads = [threading.Thread(target=client, args=(range(j, len(reqs), 3),)) for j in range(3)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        lats = [srv.wait(i) for i in ids]
        st = srv.stats()
        # a successful wait releases the request's record
        with pytest.raises(d3.D3HError) as e:
            srv.poll(ids[0])
        assert e.value.status == N.D3H_ERR_NOT_FOUND
    assert st["completed"] == len(reqs) and st["rejected"] == 0
    assert all(l > 0 for l in lats)
    for i, (p, f0) in enumerate(reqs):
        want = expected(p, f0)
        assert np.array_equal(outs[i].view(np.uint32), want.view(np.uint32)), (i, p)


def test_server_admission_by_memory_budget(ctx):
    rng = np.random.default_rng(5)
    reqs = [make_request(rng, i, lo=60, hi=64, max_steps=20) for i in range(12)]
    # room for about two requests at a time
    budget = 2 * (2 * 64 * 64 * 64 * 4) + (1 << 20)
    outs = [np.empty(f0.shape, np.float32) for _, f0 in reqs]
    with d3.Server(ctx, hbm_budget_bytes=budget, depth=2, sweep_steps=4) as srv:
        ids = [srv.submit(p, f0, d3.F32, outs[i]) for i, (p, f0) in enumerate(reqs)]
        for i in ids:
            srv.wait(i)
        # a request larger than the whole pool is rejected at submit
        big = {"nx": 256, "ny": 256, "nz": 256, "dtype": d3.F32, "steps": 1,
               "coef": d3.stable_coefficients(256, 256, 256)[0]}
        f = np.zeros((256, 256, 256), np.float32)
        with pytest.raises(d3.D3HError) as e:
            srv.submit(big, f, d3.F32, np.empty_like(f))
        assert e.value.status == 3
        assert srv.stats()["rejected"] == 1
    for i, (p, f0) in enumerate(reqs):
        assert np.array_equal(outs[i].view(np.uint32), expected(p, f0).view(np.uint32)), i


def test_newcomers_and_zero_step_requests_join_running_sweeps(ctx):
    """A large request is running when small ones, zero-step ones and a second
    large one arrive (each joins the sweeps once its own upload has completed,
    or a burst waits for its uploads): every result equals the oracle run of
    that request alone, bitwise."""
    rng = np.random.default_rng(77)
    big = []
    for i, n in enumerate((150, 120)):
        coef, _ = d3.stable_coefficients(n, n, n)
        k = d3.constants_f32(n, n, n)
        f0 = d3.add_noise_f32(d3.initialize_f32(n, n, n, k), 500 + i, 1e-3)
        big.append(({"nx": n, "ny": n, "nz": n, "dtype": d3.F32, "steps": 60, "coef": coef}, f0))
    small = [make_request(rng, i, lo=10, hi=60, max_steps=25) for i in range(12)]
    for j in (3, 7):  # zero-step requests: the result is the input
        small[j] = (dict(small[j][0], steps=0), small[j][1])
    reqs = [big[0]] + small[:6] + [big[1]] + small[6:]
    outs = [np.empty(f0.shape, np.float32) for _, f0 in reqs]
    with d3.Server(ctx, hbm_budget_bytes=1 << 30, sweep_steps=4) as srv:
        ids = [srv.submit(reqs[0][0], reqs[0][1], d3.F32, outs[0])]
        time.sleep(0.002)
        for i in range(1, len(reqs)):
            ids.append(srv.submit(reqs[i][0], reqs[i][1], d3.F32, outs[i]))
            if i % 3 == 0:
                time.sleep(0.001)
        for r in ids:
            srv.wait(r)
    for i, ((p, f0), out) in enumerate(zip(reqs, outs)):
        want = expected(p, f0)
        assert np.array_equal(out.view(np.uint32), want.view(np.uint32)), (i, p["steps"])
