#!/usr/bin/env python3
"""Benchmark of the MI355X diffusion engine; prints ONE JSON line.

Default workload: the 512^3 FP32 grid, cosine-product initial field and the
reference's coefficients, advanced by the plain drop-in kernel (one step per
launch, EXACT arithmetic, bitwise equal to the reference CPU path).  A "step"
is one diffusion time step over the whole grid.  The field is resident in HBM
before the timed region; the timed region is one synchronous d3h_field_run of
exactly --steps steps (the library records HIP events on its own stream around
the launches and synchronises on the last one, so the region is bracketed by
device synchronisation on both sides).  value = Gcell-updates/s from the host
wall clock of that call.  The steps are launched directly (--graph 0): on the
512^3 grid a hipGraph replay takes the same time per launch as direct launches
but adds a fixed cost per call (0.2-0.27 ms inside the events at 20 steps,
about 0.7 ms of wall clock at 400), against ~20 us for direct launches.

roofline: algorithmic bytes per launch (2 x storage bytes x cells, one read
and one write per cell, as the reference's GetThroughput) / average launch
duration from those events, against the 8 TB/s HBM peak; hbm_gbps_wall /
frac_wall give the same bytes over the wall-clock time per launch (the
timing of value).  traffic is read from profiles/traffic_<workload>.json
(rocprofv3 FETCH_SIZE / WRITE_SIZE), when present.

served (default workload only, --served-requests 0 skips it): the second half
of the headline metric -- the request server on a bounded sample of the
served configuration (seeded Poisson stream, dims log-uniform [32, 384],
50-2000 steps, 70 % FP32 / 30 % BF16 storage with stochastic rounding):
closed-loop solves/s, p50/p99 latency at --served-rate arrivals/s, solves/s
of the plain kernel launched per request, whether every result equals the
per-request result bitwise, and FP32 / BF16 errors against the FP64 GPU run.

temporal_blocking and batched (default workload only, --extras 0 skips them):
bounded runs of the other two benchmark configurations under the same clock --
the 768^3 FP32 field with seeded noise at depth 4 against plain steps
(--tb-steps steps each, effective Gcell/s, bitwise equality), and the batch of
512 uneven grids (200 steps) in FP32 and BF16 storage (solves/s end to end and
device-resident, against the plain kernel per request, bitwise equality,
errors against the FP64 GPU run).

error_vs_fp64_gpu_oracle: relative L2 and max-abs difference of the FP32
result from the same initial field and coefficients advanced in FP64 on the
GPU (EXACT arithmetic); error_vs_analytical_fp64: RMS / max-abs against the
exact PDE solution, reduced in FP64 on the device.  tb768_f32 starts from the
cosine product plus seeded uniform noise of amplitude 1e-3.

cpu_baseline: the reference's OpenMP RunKernel (oracle/_ref/libref.so, built
from the reference sources) or, when that library is absent, the oracle's C
restatement, timed on this host's cores on a bounded sample of the same grid.

Other workloads (--workload): plain512_f64, plain512_bf16 (BF16 storage, FP32
arithmetic, stochastic rounding unless --bf16-rounding nearest; its CPU
baseline is the FP32 reference), tb768_f32 (temporal blocking),
plain16384_2d / tb16384_2d (the diffusion/ family's 2D 5-point operator on a
16384^2 FP32 grid, 2D kernels, depth 1 / 8),
batched512_f32 / batched512_bf16 (512 uneven grids per run, solves/s against
launching the plain kernel per request) and served (request server: solves/s,
p50/p99 latency under Poisson arrivals).  Each prints its own JSON line.

Only one GPU is used: --gpus N > 1 exits with a message.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
WORKLOADS = {
    # name: (nx, ny, nz, dtype, depth, default steps, seeded noise amplitude)
    "plain512_f32": (512, 512, 512, "f32", 1, 1000, 0.0),
    "plain512_f64": (512, 512, 512, "f64", 1, 1000, 0.0),
    # BF16 storage, FP32 arithmetic (rounding per --bf16-rounding)
    "plain512_bf16": (512, 512, 512, "bf16", 1, 1000, 0.0),
    "tb768_f32": (768, 768, 768, "f32", 4, 2000, 1e-3),
    # 2D 5-point operator (nz = 1): diffusion/diffusion_openmp.cc:59-82
    "plain16384_2d": (16384, 16384, 1, "f32", 1, 1000, 0.0),
    "tb16384_2d": (16384, 16384, 1, "f32", 8, 2000, 0.0),
}
NOISE_SEED = 768
BF16_SEED = 0x5EED
# cell updates of the CPU baseline sample per --cpu-steps step (the 512^3 grid)
CPU_SAMPLE_CELLS = 512 ** 3


def cpu_baseline_2d(nx, ny, steps_sample):
    """Time the diffusion/ family's OpenMP RunKernel in 2D (oracle/_ref/libref2d.so)
    or the oracle's 2D restatement on the host."""
    gomp = ctypes.CDLL("libgomp.so.1")
    gomp.omp_get_max_threads.restype = ctypes.c_int
    threads = gomp.omp_get_max_threads()
    ref = os.path.join(ROOT, "oracle", "_ref", "libref2d.so")
    if os.path.exists(ref):
        L = ctypes.CDLL(ref)
        fp = ctypes.POINTER(ctypes.c_float)
        L.ref2d_run.argtypes = [ctypes.c_int] * 6 + [fp, fp, fp, fp, ctypes.POINTER(ctypes.c_double)]
        sec = ctypes.c_double()
        if L.ref2d_run(1, 2, nx, ny, 1, steps_sample, None, None, None, None, ctypes.byref(sec)) != 0:
            raise RuntimeError("ref2d_run failed")
        seconds, kind = sec.value, "reference"
    else:
        import oracle
        k = oracle.constants2d_f32(nx, ny)
        f = oracle.initialize2d_f32(nx, ny, k)
        g = np.empty_like(f)
        c = np.ascontiguousarray(k[8:15])
        t0 = time.perf_counter()
        oracle.lib().oracle_run2d_f32(f.ctypes.data_as(oracle._f), g.ctypes.data_as(oracle._f), nx, ny,
                                      c.ctypes.data_as(oracle._f), steps_sample)
        seconds, kind = time.perf_counter() - t0, "port"
    return {"value": nx * ny * steps_sample / seconds / 1e9, "unit": "Gcell-updates/s",
            "cores": threads, "kind": kind,
            "sample": "%dx%d FP32 2D, %d steps, OpenMP RunKernel wall clock (%.2f s)"
                      % (nx, ny, steps_sample, seconds)}


def cpu_baseline(nx, ny, nz, steps_sample, temporal_blocking=False):
    """Time the reference's OpenMP RunKernel (or the oracle port) on the host.
    temporal_blocking: the reference's Diffusion3DOpenMPTemporalBlocking
    (diffusion3d/diffusion3d_openmp_temporal_blocking.cc:8-68), which needs
    nz % threads == 0 and an even nz / threads and runs count + 1 steps for odd
    counts, so the sample is an even step count; None when the thread count
    does not divide the grid that way."""
    gomp = ctypes.CDLL("libgomp.so.1")
    gomp.omp_get_max_threads.restype = ctypes.c_int
    threads = gomp.omp_get_max_threads()
    ref = os.path.join(ROOT, "oracle", "_ref", "libref.so")
    if temporal_blocking:
        if not os.path.exists(ref) or nz % threads or (nz // threads) % 2:
            return None
        steps_sample += steps_sample % 2
    if os.path.exists(ref):
        L = ctypes.CDLL(ref)
        fp = ctypes.POINTER(ctypes.c_float)
        L.ref_run.argtypes = [ctypes.c_int] * 5 + [fp, fp, fp, fp, ctypes.POINTER(ctypes.c_double)]
        sec = ctypes.c_double()
        rc = L.ref_run(2 if temporal_blocking else 1, nx, ny, nz, steps_sample, None, None, None,
                       None, ctypes.byref(sec))
        if rc != 0:
            raise RuntimeError("ref_run failed")
        seconds, kind = sec.value, "reference"
    else:
        import oracle
        k = oracle.constants_f32(nx, ny, nz)
        f = oracle.initialize_f32(nx, ny, nz, k)
        g = np.empty_like(f)
        c = np.ascontiguousarray(k[8:15])
        L = oracle.lib()
        t0 = time.perf_counter()
        L.oracle_run_f32(f.ctypes.data_as(oracle._f), g.ctypes.data_as(oracle._f), nx, ny, nz,
                         c.ctypes.data_as(oracle._f), steps_sample)
        seconds, kind = time.perf_counter() - t0, "port"
    cells = nx * ny * nz * steps_sample
    return {"value": cells / seconds / 1e9, "unit": "Gcell-updates/s", "cores": threads,
            "kind": kind,
            "sample": "%dx%dx%d FP32, %d steps, %s RunKernel wall clock (%.2f s)"
                      % (nx, ny, nz, steps_sample,
                         "OpenMP temporal-blocking" if temporal_blocking else "OpenMP", seconds)}


def depth_desc(depth):
    return ("depth %d" % depth) if depth else "depth per size class (tuned, else FP32 2 / BF16 1)"


def fp64_errors(d3, ctx, problems, inits, outs, sample):
    """Relative L2 / max-abs of results against the FP64 GPU run of the same
    problem (exact arithmetic), over a sample of problems."""
    worst = {"rel_l2": 0.0, "max_abs": 0.0}
    for i in sample:
        p = problems[i]
        with d3.Field(ctx, p["nx"], p["ny"], p["nz"], d3.F64) as ref, \
                d3.Field(ctx, p["nx"], p["ny"], p["nz"], d3.F32) as got:
            ref.upload(inits[i], d3.F32)
            ref.run(p["steps"], p["coef"])
            got.upload(outs[i], d3.F32)
            c = got.compare(ref)
        worst["rel_l2"] = max(worst["rel_l2"], c["rel_l2"])
        worst["max_abs"] = max(worst["max_abs"], c["max_abs"])
    return worst


def bench_batched(args, d3, ctx, dname):
    from benchmarks_amd import workloads
    dtype = {"f32": d3.F32, "bf16": d3.BF16}[dname]
    # 0: per size class, the autotuner's choice or the dtype's default (2 for
    # FP32, 1 for BF16 storage)
    depth = args.depth if args.depth is not None else 0
    arith = d3.ARITH_FMA if args.arith == "fma" else d3.ARITH_EXACT
    rounding = d3.ROUND_STOCHASTIC if args.bf16_rounding == "stochastic" else d3.ROUND_NEAREST
    probs = workloads.batched_problems(dtype=dtype, rounding=rounding)
    # host fields live in pinned memory (both the batch and the per-request
    # baseline use these same buffers)
    arena = workloads.PinnedArena()
    inits, outs, base_outs = [], [], []
    for i, p in enumerate(probs):
        shape = (p["nz"], p["ny"], p["nx"])
        inits.append(workloads.initial_field(p, i, out=arena.empty(shape)))
        outs.append(arena.empty(shape))
        base_outs.append(arena.empty(shape))
    cells = sum(p["nx"] * p["ny"] * p["nz"] for p in probs)
    steps = probs[0]["steps"]
    # warm-up (kernels, allocator)
    with d3.Batch(ctx, probs[:16]) as b:
        b.upload_all(inits[:16], d3.F32)
        b.run(arith=arith, depth=depth, graph=True)
    # batch autotuner: depth, tile and z-chunks per size class, chosen on this
    # batch's own sweeps (a one-time cost; the table can be saved and reused)
    tune = None
    if args.batch_tune and depth == 0:
        t_tune = time.perf_counter()
        with d3.Batch(ctx, probs) as b:
            tune = b.autotune(arith)
        tune["seconds"] = round(time.perf_counter() - t_tune, 2)
        print("batch autotune:", tune, file=sys.stderr, flush=True)
    # end to end: create, upload, run (graph capture included), download
    t0 = time.perf_counter()
    b = d3.Batch(ctx, probs)
    b.upload_all(inits, d3.F32)
    tm = b.run(arith=arith, depth=depth, graph=True)
    b.download_all(outs, d3.F32)
    t_e2e = time.perf_counter() - t0
    # device-resident replay of the captured sequence
    b.upload_all(inits, d3.F32)
    t1 = time.perf_counter()
    tm2 = b.run(arith=arith, depth=depth, graph=True)
    t_dev = time.perf_counter() - t1
    b.close()
    print("batched: e2e %.3f s, device %.3f s (%d launches)" % (t_e2e, t_dev, tm2["launches"]),
          file=sys.stderr, flush=True)
    # baseline: the plain kernel launched per request (alloc, H2D, launches, D2H)
    t2 = time.perf_counter()
    for i, p in enumerate(probs):
        with d3.Field(ctx, p["nx"], p["ny"], p["nz"], dtype) as f:
            f.upload(inits[i], d3.F32)
            f.run(p["steps"], p["coef"], arith=arith, rounding=p.get("rounding", 0),
                  seed=p.get("seed", 0))
            f.download(d3.F32, out=base_outs[i])
    t_base = time.perf_counter() - t2
    same = all(np.array_equal(o.view(np.uint32), r.view(np.uint32)) for o, r in zip(base_outs, outs))
    err = fp64_errors(d3, ctx, probs, inits, outs, range(0, len(probs), 64))
    arena.close()
    n = len(probs)
    return {
        "metric": "batched solves/s (512 uneven grids, %d steps)" % steps,
        "value": round(n / t_e2e, 2), "unit": "solves/s", "n_gpus": 1, "steps": steps,
        "warmup": 1, "ms_per_step": round(t_e2e * 1e3 / steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dname,
        "data": "synthetic (seeded dims log-uniform [32,256], cosine product + 1e-3 noise)",
        "bf16_rounding": args.bf16_rounding if dtype == d3.BF16 else None,
        "config": {"workload": "%d grids %s, %d steps, %s, %s" % (
                       n, dname.upper(), steps,
                       "depth, tile and z-chunks per size class (batch autotuner)" if tune
                       else depth_desc(depth), args.arith),
                   "cells": cells, "host_buffers": "pinned (batch and per-request alike)",
                   "parallelism": "single GPU"},
        "device_resident_solves_per_s": round(n / t_dev, 2),
        "device_resident_gcell_updates_per_s": round(cells * steps / t_dev / 1e9, 2),
        "per_request_plain_solves_per_s": round(n / t_base, 2),
        "speedup_vs_per_request": round(t_base / t_e2e, 2),
        "results_equal_per_request": bool(same),
        "error_vs_fp64": err,
        "batch_autotune": tune,
    }


def served_measure(d3, ctx, nreq, rates, arith, depth, sweep_steps, rounding):
    """The request server on nreq seeded requests of the served configuration:
    closed loop (all submitted at once), open loop at each Poisson rate, the
    plain kernel per request (alloc, H2D, steps, D2H), bitwise comparison and
    errors against the FP64 GPU run."""
    from benchmarks_amd import workloads
    arena = workloads.PinnedArena()
    base = workloads.served_requests(nreq, lam=1.0, rounding=rounding)
    ins, outs = [], []
    for i, p in enumerate(base):
        a = arena.empty((p["nz"], p["ny"], p["nx"]))
        workloads.initial_field(p, i, out=a)
        ins.append(a)
        outs.append(arena.empty(a.shape))
    work = sum(p["nx"] * p["ny"] * p["nz"] * p["steps"] for p in base)
    srv = d3.Server(ctx, arith=arith, depth=depth, sweep_steps=sweep_steps)
    # warm-up
    w = srv.submit(base[0], ins[0], d3.F32, outs[0])
    srv.wait(w)
    # closed loop: everything submitted at once
    t0 = time.perf_counter()
    ids = [srv.submit(p, ins[i], d3.F32, outs[i]) for i, p in enumerate(base)]
    for r in ids:
        srv.wait(r)
    t_closed = time.perf_counter() - t0
    server_outs = [o.copy() for o in outs]
    print("served closed loop: %d requests in %.3f s" % (nreq, t_closed), file=sys.stderr, flush=True)
    # open loop at each arrival rate
    rates_out = []
    for lam in rates:
        reqs = workloads.served_requests(nreq, lam=lam, rounding=rounding)
        lat = []
        t_start = time.perf_counter()
        ids = []
        for i, p in enumerate(reqs):
            delay = p["arrival_s"] - (time.perf_counter() - t_start)
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(p, ins[i], d3.F32, outs[i]))
        for r in ids:
            lat.append(srv.wait(r))
        t_open = time.perf_counter() - t_start
        lat = np.array(lat)
        rates_out.append({"rate_per_s": lam, "solves_per_s": round(nreq / t_open, 2),
                          "p50_ms": round(float(np.percentile(lat, 50)), 2),
                          "p99_ms": round(float(np.percentile(lat, 99)), 2)})
        print("rate %.0f/s: %s" % (lam, rates_out[-1]), file=sys.stderr, flush=True)
    st = srv.stats()
    srv.close()
    # baseline: the plain kernel launched per request, one after another
    t1 = time.perf_counter()
    same = True
    for i, p in enumerate(base):
        with d3.Field(ctx, p["nx"], p["ny"], p["nz"], p["dtype"]) as f:
            f.upload(ins[i], d3.F32)
            f.run(p["steps"], p["coef"], arith=arith, rounding=p.get("rounding", 0),
                  seed=p.get("seed", 0))
            o = f.download(d3.F32)
        same = same and np.array_equal(o.view(np.uint32), server_outs[i].view(np.uint32))
    t_base = time.perf_counter() - t1
    f32 = [i for i, p in enumerate(base) if p["dtype"] == d3.F32][:6]
    bf = [i for i, p in enumerate(base) if p["dtype"] == d3.BF16][:6]
    e32 = fp64_errors(d3, ctx, base, ins, server_outs, f32)
    ebf = fp64_errors(d3, ctx, base, ins, server_outs, bf)
    arena.close()
    return {"requests": nreq, "cell_updates": work, "t_closed": t_closed, "t_base": t_base,
            "open_loop": rates_out, "results_equal_per_request": bool(same), "server_stats": st,
            "error_vs_fp64": {"f32": e32, "bf16": ebf}}


def served_summary(d3, ctx, args):
    """Bounded served figures for the default (headline) line."""
    arith = d3.ARITH_FMA if args.arith == "fma" else d3.ARITH_EXACT
    r = served_measure(d3, ctx, args.served_requests, [args.served_rate], arith, 0,
                       args.sweep_steps, d3.ROUND_STOCHASTIC)
    n = r["requests"]
    return {
        "workload": "%d seeded requests: edge log-uniform [32,384], slab aspect, steps U[50,2000], "
                    "70%% FP32 / 30%% BF16 storage (stochastic rounding), depth per size class, "
                    "%d steps per sweep, %s arithmetic" % (n, args.sweep_steps, args.arith),
        "closed_loop_solves_per_s": round(n / r["t_closed"], 2),
        "closed_loop_gcell_updates_per_s": round(r["cell_updates"] / r["t_closed"] / 1e9, 2),
        "per_request_plain_solves_per_s": round(n / r["t_base"], 2),
        "speedup_vs_per_request": round(r["t_base"] / r["t_closed"], 2),
        "open_loop": r["open_loop"][0],
        "results_equal_per_request": r["results_equal_per_request"],
        "error_vs_fp64": r["error_vs_fp64"],
    }


def tb_summary(d3, ctx, steps):
    """Bounded temporal-blocking figures for the default line: the 768^3 FP32
    field with seeded noise, depth 4 (autotuned tile), `steps` steps from the
    same resident input as `steps` plain steps, whose result it must equal
    bitwise."""
    n = 768
    k = d3.constants_f32(n, n, n)
    f0 = d3.add_noise_f32(d3.initialize_f32(n, n, n, k), NOISE_SEED, 1e-3)
    coef = d3.coefficients(k)
    out = {"workload": "768^3 FP32 + seeded noise 1e-3, %d steps, depth 4 vs plain, exact "
                       "arithmetic" % steps}
    with d3.Field(ctx, n, n, n, d3.F32) as pristine, d3.Field(ctx, n, n, n, d3.F32) as f:
        pristine.upload(f0)
        del f0
        ctx.autotune(n, n, n, d3.F32, 4, d3.ARITH_EXACT)
        ctx.autotune(n, n, n, d3.F32, 1, d3.ARITH_EXACT)
        res = {}
        for depth in (4, 1):
            f.copy_from(pristine)
            f.run(8, coef, depth=depth)                 # warm-up
            f.copy_from(pristine)
            t0 = time.perf_counter()
            tm = f.run(steps, coef, depth=depth)
            wall = time.perf_counter() - t0
            res[depth] = f.download()
            cells = n ** 3
            out["depth%d" % depth] = {
                "gcell_updates_per_s": round(cells * steps / wall / 1e9, 2),
                "avg_launch_ms": round(tm["avg_launch_ms"], 4),
                "effective_frac_of_hbm_peak": round(
                    8.0 * cells * tm["temporal_depth"] / (tm["avg_launch_ms"] * 1e-3) / 1e9
                    / HBM_PEAK_GBPS, 4),
                "tile": tm["tile"], "zchunks": tm["zchunks"], "pace_ns": tm["pace_ns"],
                "launches": tm["launches"]}
    out["speedup_depth4_vs_plain"] = round(out["depth4"]["gcell_updates_per_s"]
                                           / out["depth1"]["gcell_updates_per_s"], 3)
    out["bitwise_equal_to_plain"] = bool(np.array_equal(res[4].view(np.uint32),
                                                        res[1].view(np.uint32)))
    return out


def batched_summary(args, d3, ctx):
    """Bounded batched figures for the default line: the 512-grid batch (200
    steps) in FP32 and in BF16 storage with stochastic rounding."""
    out = {}
    for dname in ("f32", "bf16"):
        r = bench_batched(args, d3, ctx, dname)
        out[dname] = {k: r[k] for k in (
            "value", "device_resident_solves_per_s", "device_resident_gcell_updates_per_s",
            "per_request_plain_solves_per_s", "speedup_vs_per_request",
            "results_equal_per_request", "error_vs_fp64")}
        out[dname]["solves_per_s"] = out[dname].pop("value")
    out["workload"] = ("512 uneven grids (dims log-uniform [32,256], seeded), 200 steps, depth per "
                       "size class; solves/s end to end (upload, run, download) vs the plain "
                       "kernel launched per request")
    return out


def bench_served(args, d3, ctx):
    arith = d3.ARITH_FMA if args.arith == "fma" else d3.ARITH_EXACT
    depth = args.depth if args.depth is not None else 0
    nreq = args.requests
    rounding = d3.ROUND_STOCHASTIC if args.bf16_rounding == "stochastic" else d3.ROUND_NEAREST
    r = served_measure(d3, ctx, nreq, [float(x) for x in args.rates.split(",")], arith, depth,
                       args.sweep_steps, rounding)
    t_closed, t_base, work = r["t_closed"], r["t_base"], r["cell_updates"]
    return {
        "metric": "served solves/s, p50/p99 latency at equal error",
        "value": round(nreq / t_closed, 2), "unit": "solves/s", "n_gpus": 1, "steps": nreq,
        "warmup": 1, "ms_per_step": round(t_closed * 1e3 / nreq, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32/bf16 storage, f32 compute",
        "data": "synthetic (seeded requests: edge log-uniform [32,384], slab aspect, steps U[50,2000], "
                "70% FP32 / 30% BF16)",
        "bf16_rounding": args.bf16_rounding,
        "config": {"workload": "%d requests, %s, %d steps per sweep, %s" % (
                       nreq, depth_desc(depth), args.sweep_steps, args.arith),
                   "cell_updates": work, "parallelism": "single GPU"},
        "closed_loop_gcell_updates_per_s": round(work / t_closed / 1e9, 2),
        "per_request_plain_solves_per_s": round(nreq / t_base, 2),
        "speedup_vs_per_request": round(t_base / t_closed, 2),
        "results_equal_per_request": r["results_equal_per_request"],
        "open_loop": r["open_loop"],
        "server_stats": r["server_stats"],
        "error_vs_fp64": r["error_vs_fp64"],
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="plain512_f32",
                    choices=sorted(WORKLOADS) + ["batched512_f32", "batched512_bf16", "served"])
    ap.add_argument("--arith", default="exact", choices=["exact", "fma"])
    ap.add_argument("--depth", type=int, default=None,
                    help="temporal depth (default per workload); 0 = the autotuner's choice")
    ap.add_argument("--graph", type=int, default=0,
                    help="replay a captured hipGraph (1) or launch the steps directly (0, default)")
    ap.add_argument("--tune", type=int, default=1, help="autotune the tile before timing")
    ap.add_argument("--batch-tune", type=int, default=0,
                    help="batched: run the batch autotuner (depth, tile, z-chunks per size class) "
                         "before timing")
    ap.add_argument("--cpu-steps", type=int, default=500, help="steps of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 GPU oracle comparison")
    ap.add_argument("--bf16-rounding", default="stochastic", choices=["stochastic", "nearest"],
                    help="rounding of BF16 storage in the batched and served workloads")
    ap.add_argument("--requests", type=int, default=150, help="served: requests per run")
    ap.add_argument("--rates", default="50,100,150", help="served: Poisson arrival rates (1/s)")
    ap.add_argument("--sweep-steps", type=int, default=32, help="served: steps per scheduling round")
    ap.add_argument("--served-requests", type=int, default=150,
                    help="default workload: requests of the bounded served sample (0 = skip)")
    ap.add_argument("--served-rate", type=float, default=50.0,
                    help="default workload: Poisson arrival rate (1/s) of the served sample's "
                         "open-loop run")
    ap.add_argument("--extras", type=int, default=1,
                    help="default workload: add the bounded depth-4 768^3 and batched 512-grid "
                         "objects to the line (0 = skip)")
    ap.add_argument("--tb-steps", type=int, default=200,
                    help="default workload: steps of the depth-4 768^3 comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU (this build has no multi-GPU path); "
              "got --gpus %d" % args.gpus, file=sys.stderr)
        return 2

    import benchmarks_amd as d3
    if args.workload.startswith("batched") or args.workload == "served":
        ctx = d3.Context(0)
        if args.workload == "served":
            line = bench_served(args, d3, ctx)
        else:
            line = bench_batched(args, d3, ctx, args.workload.split("_")[1])
        print(json.dumps(line), flush=True)
        ctx.close()
        return 0
    nx, ny, nz, dname, depth, dsteps, noise = WORKLOADS[args.workload]
    steps = args.steps if args.steps is not None else dsteps
    depth = args.depth if args.depth is not None else depth
    dtype = {"f32": d3.F32, "f64": d3.F64, "bf16": d3.BF16}[dname]
    arith = d3.ARITH_FMA if args.arith == "fma" else d3.ARITH_EXACT
    esize = {d3.F32: 4, d3.F64: 8, d3.BF16: 2}[dtype]
    # BF16 storage rounds every stored level (stochastically with a fixed seed
    # unless --bf16-rounding nearest)
    rnd = {}
    if dtype == d3.BF16 and args.bf16_rounding == "stochastic":
        rnd = {"rounding": d3.ROUND_STOCHASTIC, "seed": BF16_SEED}

    ctx = d3.Context(0)
    info = ctx.info()
    print("device: %d CUs, %.1f GB HBM" % (info["num_cus"], info["total_bytes"] / 1e9),
          file=sys.stderr, flush=True)
    two_d = nz == 1
    if two_d:
        # the diffusion/ family's 2D benchmark: rows reversed on the device so
        # the y-1 term is summed first, ct = cb = 0
        k = d3.constants2d_f32(nx, ny)

        def initial():
            return d3.initialize2d_f32(nx, ny, k)
    elif dtype in (d3.F32, d3.BF16):
        k = d3.constants_f32(nx, ny, nz)

        def initial():
            f = d3.initialize_f32(nx, ny, nz, k)
            return d3.add_noise_f32(f, NOISE_SEED, noise) if noise else f
    else:
        k = d3.constants_f64(nx, ny, nz)

        def initial():
            return d3.initialize_f64(nx, ny, nz, k)
    coef = d3.coefficients(k)
    field = d3.Field(ctx, nx, ny, nz, dtype)
    # the initial field stays on the device, so the input is restored right
    # after the warm-up by a device copy (no host round trip and no idle gap
    # before the timed region)
    pristine = d3.Field(ctx, nx, ny, nz, dtype)
    if two_d:
        field.set_y_reversed(True)
        pristine.set_y_reversed(True)
    pristine.upload(initial())
    # the probe and the autotuner run after the host-side initialisation, so
    # the device stays busy from them through the warm-up to the timed run
    probe = ctx.bandwidth_probe(1 << 28, d3.F32, reps=10)
    print("copy probe: %.1f GB/s" % probe["gbps"], file=sys.stderr, flush=True)
    tuned = None
    if args.tune:
        tuned = ctx.autotune(nx, ny, nz, dtype, depth, arith)
        print("tuned:", tuned, file=sys.stderr, flush=True)

    field.copy_from(pristine)
    if args.graph:
        # capture once (untimed), then restore the input
        field.run(steps, coef, depth=depth, arith=arith, graph=True, **rnd)
        field.copy_from(pristine)
    # The timed field is re-initialised before the warm-up, which runs on the
    # other field of the same shape (the reference warms up and re-initialises,
    # diffusion/diffusion.h:126-131): the launches right before the timed
    # region are then the stencil's own, not the restoring device copy, after
    # which the first launches run 15-30 us slower.
    if args.warmup > 0:
        pristine.run(args.warmup, coef, depth=depth, arith=arith, graph=False, **rnd)
    t0 = time.perf_counter()
    tm = field.run(steps, coef, depth=depth, arith=arith, graph=bool(args.graph), **rnd)
    wall = time.perf_counter() - t0
    print("timing:", tm, file=sys.stderr, flush=True)

    cells = nx * ny * nz
    gcells = cells * steps / wall / 1e9
    wave = [float(v) for v in k[4:7]]
    if two_d:
        # no z factor in 2D: kz = pi with dz = 1 gives cos(kz * 0.5) = 0, factor 1
        wave[2] = float(np.pi)
    err = field.error_analytical(float(k[0]), float(steps) * float(k[7]), wave,
                                 [float(v) for v in k[1:4]])
    # the FP64 GPU oracle: the same initial field and coefficients advanced in
    # FP64 (EXACT arithmetic; fused depths are bitwise equal to plain steps)
    err64 = None
    if dtype in (d3.F32, d3.BF16) and not args.no_fp64:
        t64 = time.perf_counter()
        with d3.Field(ctx, nx, ny, nz, d3.F64) as ref:
            if two_d:
                ref.set_y_reversed(True)
            ref.upload(initial())
            ref.run(steps, coef, depth=1 if two_d else 4)
            err64 = field.compare(ref)
        print("fp64 oracle: %.1f s" % (time.perf_counter() - t64), file=sys.stderr, flush=True)
    field.close()
    pristine.close()

    bytes_per_launch = 2.0 * esize * cells          # one read + one write per cell
    avg_launch_s = tm["avg_launch_ms"] * 1e-3
    steps_per_launch = tm["temporal_depth"] if tm["launches"] == steps // tm["temporal_depth"] else 1
    # for fused launches the algorithmic bytes are those of the single step
    # sequence they replace (effective bandwidth)
    achieved = bytes_per_launch * steps_per_launch / avg_launch_s / 1e9
    # the same bytes over the wall-clock time per launch (the timing of value)
    achieved_wall = bytes_per_launch * steps_per_launch * tm["launches"] / wall / 1e9
    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated
    # against the copy probe) when the profile is of the kernel that ran here
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.workload)
    if os.path.exists(tpath):
        with open(tpath) as fh:
            prof = json.load(fh)
        t = d3.tiles()[tm["tile"]]
        vec = t["tile_x"] // (64 * t["groups"])
        fma = "true" if t["arith"] else "false"
        if t["row_stream"]:
            name = "stencil2d_kernel<%d, %d, %d, %d, %d, %s, %d>" % (
                t["dtype"], vec, t["temporal_depth"], t["groups"], t["rows_per_wave"], fma,
                t["prefetch"])
        else:
            name = "stencil_kernel<%d, %d, %d, %d, %d, %d, %s, %d>" % (
                t["dtype"], vec, t["temporal_depth"], t["groups"], t["rows_per_wave"], t["waves"],
                fma, t["prefetch"])
        grid = tm["blocks"] * 64 * t["waves"]
        # one profiled kernel, or a table of every kernel and grid the
        # autotuner launched in the profiled runs ("dispatches")
        for d in prof.get("dispatches", [prof]):
            if d.get("kernel") == name and d.get("grid_size", grid) == grid:
                traffic = d.get("hbm_bytes_per_launch")
                break

    depth_used = tm["temporal_depth"]
    cpu_plain = None
    served = None
    if args.workload == "plain512_f32" and args.served_requests > 0:
        served = served_summary(d3, ctx, args)
        print("served:", served, file=sys.stderr, flush=True)
    tb_line = batched_line = None
    if args.workload == "plain512_f32" and args.extras:
        tb_line = tb_summary(d3, ctx, args.tb_steps)
        print("temporal blocking:", tb_line, file=sys.stderr, flush=True)
        batched_line = batched_summary(args, d3, ctx)
        print("batched:", batched_line, file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        if two_d:
            cpu = cpu_baseline_2d(nx, ny, max(1, args.cpu_steps * CPU_SAMPLE_CELLS // (nx * ny)))
        else:
            # a sample of about --cpu-steps 512^3 steps' worth of cell updates
            sample = max(2, args.cpu_steps * CPU_SAMPLE_CELLS // (nx * ny * nz))
            cpu = cpu_baseline(nx, ny, nz, sample)
            if depth_used > 1:
                # the reference's own CPU temporal blocking beside its plain
                # OpenMP path (the GPU counterpart of the fused launches)
                # (it runs several times slower than the plain OpenMP path:
                # per-iteration work-slab allocation and the extra planes, so
                # it gets a quarter of the sample)
                tb = cpu_baseline(nx, ny, nz, max(2, sample // 4), temporal_blocking=True)
                if tb is not None:
                    cpu, cpu_plain = tb, cpu

    size = ("%d^2 (2D 5-point)" % nx) if two_d else ("%d^3" % nx)
    if dtype == d3.BF16:
        size += " (BF16 storage)"
    metric = "Gcell-updates/s + %%HBM roofline at %s" % size
    if depth_used > 1:
        metric = "effective Gcell-updates/s + %%HBM roofline at %s, temporal depth %d" % (size,
                                                                                       depth_used)
    line = {
        "metric": metric,
        "value": round(gcells, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 storage, f32 compute" if dtype == d3.BF16 else dname,
        "bf16_rounding": args.bf16_rounding if dtype == d3.BF16 else None,
        "data": "synthetic (analytical cosine-product field%s, reference coefficients)"
                % (", + seeded uniform noise amplitude %g" % noise if noise else ""),
        "config": {"workload": "%dx%dx%d %s %s diffusion, %d steps, depth %d, %s arithmetic"
                               % (nx, ny, nz, dname.upper(), "2D 5-point" if two_d else "7-point",
                                  steps, depth_used, args.arith),
                   "grid": [nx, ny, nz], "temporal_depth": depth_used,
                   "depth_choice": "autotuned" if depth == 0 else "fixed", "arith": args.arith,
                   "graph": bool(args.graph), "tile": tm["tile"], "zchunks": tm["zchunks"],
                   "pace_ns": tm["pace_ns"], "blocks": tm["blocks"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": traffic},
        "cpu_baseline": cpu,
        "cpu_baseline_plain_openmp": cpu_plain,
        "kernel_ms": round(tm["kernel_ms"], 3),
        "avg_launch_ms": round(tm["avg_launch_ms"], 5),
        "hbm_gbps_wall": round(achieved_wall, 1),
        "frac_wall": round(achieved_wall / HBM_PEAK_GBPS, 4),
        "copy_probe_gbps": round(probe["gbps"], 1),
        "frac_of_copy_probe": round(achieved / probe["gbps"], 4),
        "error_vs_analytical_fp64": {"rms": err["rms"], "max_abs": err["max_abs"]},
        "error_vs_fp64_gpu_oracle": ({"rel_l2": err64["rel_l2"], "max_abs": err64["max_abs"]}
                                     if err64 else None),
        "served": served,
        "temporal_blocking": tb_line,
        "batched": batched_line,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
