"""Python host API over the C ABI: contexts, device-resident fields, batches,
servers and the host-side analytical field.

Arrays cross the boundary as dense numpy arrays in (z, y, x) order, i.e. the
reference's OFFSET3D layout (x fastest, diffusion3d/diffusion3d.h:33-34).
"""
import ctypes

import numpy as np

from . import _native as N
from ._native import check, lib

NP_DTYPES = {N.F32: np.float32, N.F64: np.float64, N.BF16: np.uint16, N.F16: np.float16}
DTYPE_NAMES = {N.F32: "f32", N.F64: "f64", N.BF16: "bf16", N.F16: "f16"}
DTYPE_BY_NAME = {v: k for k, v in DTYPE_NAMES.items()}


def _dptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _coef_array(coef):
    c = (ctypes.c_double * 7)(*[float(v) for v in coef])
    return c


# ---- host-side benchmark constants and analytical field ----------------------

def constants_f32(nx, ny, nz):
    """kappa, dx, dy, dz, kx, ky, kz, dt, ce, cw, cn, cs, ct, cb, cc as float32
    (diffusion3d/diffusion3d.h:54-66)."""
    out = np.zeros(15, dtype=np.float32)
    check(lib().d3h_constants_f32(nx, ny, nz, out.ctypes.data_as(N.P(ctypes.c_float))),
          "d3h_constants_f32")
    return out


def constants_f64(nx, ny, nz):
    out = np.zeros(15, dtype=np.float64)
    check(lib().d3h_constants_f64(nx, ny, nz, out.ctypes.data_as(N.P(ctypes.c_double))),
          "d3h_constants_f64")
    return out


def coefficients(constants):
    """The 7 stencil coefficients ce cw cn cs ct cb cc of a constants vector."""
    return [float(v) for v in constants[8:15]]


def initialize_f32(nx, ny, nz, constants, time=0.0, out=None):
    """Cosine-product field at `time` (diffusion3d/diffusion3d.cc:14-39)."""
    if out is None:
        out = np.empty((nz, ny, nx), dtype=np.float32)
    k = np.ascontiguousarray(constants, dtype=np.float32)
    check(lib().d3h_initialize_f32(out.ctypes.data_as(N.P(ctypes.c_float)), nx, ny, nz,
                                   k.ctypes.data_as(N.P(ctypes.c_float)), float(time)),
          "d3h_initialize_f32")
    return out


def initialize_f64(nx, ny, nz, constants, time=0.0, out=None):
    if out is None:
        out = np.empty((nz, ny, nx), dtype=np.float64)
    k = np.ascontiguousarray(constants, dtype=np.float64)
    check(lib().d3h_initialize_f64(out.ctypes.data_as(N.P(ctypes.c_double)), nx, ny, nz,
                                   k.ctypes.data_as(N.P(ctypes.c_double)), float(time)),
          "d3h_initialize_f64")
    return out


def constants2d_f32(nx, ny):
    """Constants of the diffusion/ family's 2D benchmark, Diffusion(2, {nx, ny})
    (diffusion/diffusion.h:93-108), in the same 15-slot order; ct = cb = 0."""
    out = np.zeros(15, dtype=np.float32)
    check(lib().d3h_constants2d_f32(nx, ny, out.ctypes.data_as(N.P(ctypes.c_float))),
          "d3h_constants2d_f32")
    return out


def initialize2d_f32(nx, ny, constants, time=0.0, out=None):
    """2D cosine product (diffusion/diffusion.h:48-77, ndim = 2) as (ny, nx)."""
    if out is None:
        out = np.empty((ny, nx), dtype=np.float32)
    k = np.ascontiguousarray(constants, dtype=np.float32)
    check(lib().d3h_initialize2d_f32(out.ctypes.data_as(N.P(ctypes.c_float)), nx, ny,
                                     k.ctypes.data_as(N.P(ctypes.c_float)), float(time)),
          "d3h_initialize2d_f32")
    return out


def save_raw(path, array, dtype=None, steps=0, y_reversed=False):
    """Write a dense (nz, ny, nx) host array as a raw field file (bit-exact)."""
    if dtype is None:
        dtype = {np.dtype(np.float32): N.F32, np.dtype(np.float64): N.F64,
                 np.dtype(np.float16): N.F16, np.dtype(np.uint16): N.BF16}[array.dtype]
    a = np.ascontiguousarray(array)
    if a.ndim != 3 or a.dtype != np.dtype(NP_DTYPES[dtype]):
        raise ValueError("save_raw needs a (nz, ny, nx) array of the given dtype")
    nz, ny, nx = a.shape
    check(lib().d3h_raw_write(str(path).encode(), _dptr(a), dtype, nx, ny, nz, steps,
                              1 if y_reversed else 0), "d3h_raw_write")


def load_raw(path):
    """Read a raw field file: returns (array (nz, ny, nx), info dict)."""
    dt, nx, ny, nz, fl = (ctypes.c_int() for _ in range(5))
    st = ctypes.c_uint64()
    p = str(path).encode()
    check(lib().d3h_raw_read(p, None, 0, ctypes.byref(dt), ctypes.byref(nx), ctypes.byref(ny),
                             ctypes.byref(nz), ctypes.byref(st), ctypes.byref(fl)), "d3h_raw_read")
    out = np.empty((nz.value, ny.value, nx.value), dtype=NP_DTYPES[dt.value])
    check(lib().d3h_raw_read(p, _dptr(out), out.nbytes, None, None, None, None, None, None),
          "d3h_raw_read")
    return out, {"dtype": dt.value, "steps": st.value, "y_reversed": bool(fl.value & 1)}


def add_noise_f32(buf, seed, amp):
    check(lib().d3h_add_noise_f32(buf.ctypes.data_as(N.P(ctypes.c_float)), buf.size, seed,
                                  float(amp)), "d3h_add_noise_f32")
    return buf


def tiles():
    """The compiled kernel/tile table."""
    out = []
    for i in range(lib().d3h_tile_count()):
        d = N.TileDesc()
        check(lib().d3h_tile_info(i, ctypes.byref(d)), "d3h_tile_info")
        out.append({k: getattr(d, k) for k, _ in d._fields_} | {"index": i})
    return out


class Context:
    """One device context (compute + copy streams, events, tuned tile table)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().d3h_context_create(device, ctypes.byref(h)), "d3h_context_create")
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            lib().d3h_context_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def info(self):
        cus, fr, tot, lds = ctypes.c_int(), ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_int()
        check(lib().d3h_device_info(self.handle, ctypes.byref(cus), ctypes.byref(fr),
                                    ctypes.byref(tot), ctypes.byref(lds)), "d3h_device_info")
        return {"num_cus": cus.value, "free_bytes": fr.value, "total_bytes": tot.value,
                "lds_bytes": lds.value}

    def field(self, nx, ny, nz, dtype=N.F32):
        return Field(self, nx, ny, nz, dtype)

    def autotune(self, nx, ny, nz, dtype=N.F32, depth=1, arith=N.ARITH_EXACT):
        t, z, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(lib().d3h_autotune(self.handle, nx, ny, nz, dtype, depth, arith, ctypes.byref(t),
                                 ctypes.byref(z), ctypes.byref(ms)), "d3h_autotune")
        return {"tile": t.value, "zchunks": z.value, "ms_per_step": ms.value}

    def save_tuning(self, path):
        check(lib().d3h_autotune_save(self.handle, path.encode()), "d3h_autotune_save")

    def load_tuning(self, path):
        check(lib().d3h_autotune_load(self.handle, path.encode()), "d3h_autotune_load")

    def bandwidth_probe(self, n, dtype=N.F32, reps=20):
        g, ms = ctypes.c_double(), ctypes.c_double()
        check(lib().d3h_bandwidth_probe(self.handle, n, dtype, reps, ctypes.byref(g),
                                        ctypes.byref(ms)), "d3h_bandwidth_probe")
        return {"gbps": g.value, "avg_ms": ms.value}


class Field:
    """A device-resident nx*ny*nz field (two ping-pong buffers)."""

    def __init__(self, ctx, nx, ny, nz, dtype=N.F32):
        h = ctypes.c_void_p()
        check(lib().d3h_field_create(ctx.handle, nx, ny, nz, dtype, ctypes.byref(h)),
              "d3h_field_create")
        self.ctx, self.handle = ctx, h
        self.nx, self.ny, self.nz, self.dtype = nx, ny, nz, dtype

    def close(self):
        if self.handle:
            lib().d3h_field_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, host, host_dtype=None):
        if host_dtype is None:
            host_dtype = {np.dtype(np.float32): N.F32, np.dtype(np.float64): N.F64,
                          np.dtype(np.float16): N.F16, np.dtype(np.uint16): N.BF16}[host.dtype]
        host = np.ascontiguousarray(host)
        if host.size != self.nx * self.ny * self.nz:
            raise ValueError("host array has %d cells, field %d" % (host.size,
                                                                     self.nx * self.ny * self.nz))
        if host.dtype != np.dtype(NP_DTYPES[host_dtype]):
            raise ValueError("host array dtype %s does not match %s" % (host.dtype,
                                                                         DTYPE_NAMES[host_dtype]))
        check(lib().d3h_field_upload(self.handle, _dptr(host), host_dtype), "d3h_field_upload")

    def copy_from(self, other):
        """Device-to-device copy of `other`'s current contents (same dims and
        dtype) into this field, with its step count."""
        check(lib().d3h_field_copy(self.handle, other.handle), "d3h_field_copy")

    def save(self, path):
        """Bit-exact raw field file (header + dense data in the field's dtype)."""
        check(lib().d3h_field_save(self.handle, str(path).encode()), "d3h_field_save")

    def load(self, path):
        """Restore a field saved by save() (same dimensions)."""
        check(lib().d3h_field_load(self.handle, str(path).encode()), "d3h_field_load")

    def set_y_reversed(self, reversed_=True):
        """Keep the rows of every plane reversed on the device so the y-1 term
        is summed before the y+1 term (the diffusion/ family's order)."""
        check(lib().d3h_field_set_y_reversed(self.handle, 1 if reversed_ else 0),
              "d3h_field_set_y_reversed")

    def download(self, host_dtype=None, out=None):
        if host_dtype is None:
            host_dtype = self.dtype
        if out is None:
            out = np.empty((self.nz, self.ny, self.nx), dtype=NP_DTYPES[host_dtype])
        elif (not out.flags.c_contiguous or out.dtype != np.dtype(NP_DTYPES[host_dtype])
              or out.size != self.nx * self.ny * self.nz):
            raise ValueError("out must be a dense array of the field's size and host dtype")
        check(lib().d3h_field_download(self.handle, _dptr(out), host_dtype), "d3h_field_download")
        return out

    def run(self, count, coef, depth=1, arith=N.ARITH_EXACT, tile=-1, zchunks=0, graph=False,
            rounding=N.ROUND_NEAREST, seed=0, pace_ns=0):
        """Advance `count` steps.  rounding=ROUND_STOCHASTIC (BF16 fields) rounds
        every stored level stochastically with `seed`; step numbers continue
        from the last upload.  pace_ns: plane schedule of depth-1 launches
        (0 = the tuned one, -1 = none, > 0 = ns per z iteration)."""
        cfg = N.RunConfig(depth, arith, tile, zchunks, 1 if graph else 0, rounding, seed, pace_ns)
        t = N.Timing()
        check(lib().d3h_field_run(self.handle, count, _coef_array(coef), ctypes.byref(cfg),
                                  ctypes.byref(t)), "d3h_field_run")
        return t.as_dict()

    def compare(self, other):
        out = (ctypes.c_double * 3)()
        check(lib().d3h_field_compare(self.handle, other.handle, out), "d3h_field_compare")
        return {"rel_l2": out[0], "max_abs": out[1], "l2": out[2]}

    def error_analytical(self, kappa, time, k, d):
        out = (ctypes.c_double * 2)()
        kk = (ctypes.c_double * 3)(*k)
        dd = (ctypes.c_double * 3)(*d)
        check(lib().d3h_field_error_analytical(self.handle, kappa, time, kk, dd, out),
              "d3h_field_error_analytical")
        return {"rms": out[0], "max_abs": out[1]}

    def error_analytical_noalloc(self, kappa, time, k, d):
        """error_analytical without any allocation (a static device scratch
        shared by every call on the device: calls must not overlap)."""
        out = (ctypes.c_double * 2)()
        kk = (ctypes.c_double * 3)(*k)
        dd = (ctypes.c_double * 3)(*d)
        check(lib().d3h_field_error_analytical_noalloc(self.handle, kappa, time, kk, dd, out),
              "d3h_field_error_analytical_noalloc")
        return {"rms": out[0], "max_abs": out[1]}


def run_kernel2d(count, f1, coef):
    """The diffusion/ family's 2D RunKernel(count) (diffusion/diffusion_openmp.cc:59-82)
    on a float32 (ny, nx) array, in place: an nx x ny x 1 device field with
    reversed rows and ct = cb = 0, which the library advances with its 2D
    5-point kernel (stencil2d_kernel)."""
    ny, nx = f1.shape
    c = [float(v) for v in coef]
    c[4] = c[5] = 0.0
    with Context() as ctx, Field(ctx, nx, ny, 1) as f:
        f.set_y_reversed(True)
        f.upload(f1)
        f.run(count, c)
        f1[...] = f.download().reshape(ny, nx)
    return f1


def run_kernel(count, f1, coef):
    """The Physis-shaped drop-in sequence on a float32 (nz, ny, nx) array, in place:
    initialize_hip / initialize_benchmark_hip / run_kernel_hip / finalize_benchmark_hip.
    Returns the kernel-only seconds."""
    nz, ny, nx = f1.shape
    L = lib()
    check(L.initialize_hip(0, None, nx, ny, nz), "initialize_hip")
    try:
        check(L.initialize_benchmark_hip(nx, ny, nz), "initialize_benchmark_hip")
        check(L.run_kernel_hip(count, f1.ctypes.data_as(N.P(ctypes.c_float)), nx, ny, nz,
                               *[float(c) for c in coef]), "run_kernel_hip")
        s = ctypes.c_float()
        check(L.kernel_time_hip(ctypes.byref(s)), "kernel_time_hip")
        return s.value
    finally:
        L.finalize_benchmark_hip()


class Batch:
    """Many independent grids advanced together (d3h_batch_*).

    problems: list of dicts with nx, ny, nz, dtype, steps, coef (7 values) and
    optionally rounding (ROUND_NEAREST / ROUND_STOCHASTIC) and seed."""

    def __init__(self, ctx, problems):
        arr = (N.Problem * len(problems))()
        for i, p in enumerate(problems):
            arr[i].nx, arr[i].ny, arr[i].nz = p["nx"], p["ny"], p["nz"]
            arr[i].dtype, arr[i].steps = p.get("dtype", N.F32), p["steps"]
            for j, c in enumerate(p["coef"]):
                arr[i].coef[j] = float(c)
            arr[i].rounding = p.get("rounding", N.ROUND_NEAREST)
            arr[i].seed = p.get("seed", 0)
        h = ctypes.c_void_p()
        check(lib().d3h_batch_create(ctx.handle, arr, len(problems), ctypes.byref(h)),
              "d3h_batch_create")
        self.ctx, self.handle, self.problems = ctx, h, list(problems)

    def close(self):
        if self.handle:
            lib().d3h_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, i, host, host_dtype):
        host = np.ascontiguousarray(host)
        check(lib().d3h_batch_upload(self.handle, i, _dptr(host), host_dtype), "d3h_batch_upload")

    def download(self, i, host_dtype=None):
        p = self.problems[i]
        if host_dtype is None:
            host_dtype = p.get("dtype", N.F32)
        out = np.empty((p["nz"], p["ny"], p["nx"]), dtype=NP_DTYPES[host_dtype])
        check(lib().d3h_batch_download(self.handle, i, _dptr(out), host_dtype),
              "d3h_batch_download")
        return out

    def _ptrs(self, arrays, host_dtype):
        if len(arrays) != len(self.problems):
            raise ValueError("%d arrays for %d problems" % (len(arrays), len(self.problems)))
        for a, p in zip(arrays, self.problems):
            if not a.flags.c_contiguous or a.dtype != np.dtype(NP_DTYPES[host_dtype]) or \
                    a.size != p["nx"] * p["ny"] * p["nz"]:
                raise ValueError("host arrays must be dense, of the problem's size and host dtype")
        return (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])

    def upload_all(self, hosts, host_dtype):
        """Upload every problem's initial field in one call (one synchronisation)."""
        check(lib().d3h_batch_upload_all(self.handle, self._ptrs(hosts, host_dtype), host_dtype),
              "d3h_batch_upload_all")

    def download_all(self, outs, host_dtype):
        """Download every problem into the given dense arrays (one synchronisation)."""
        check(lib().d3h_batch_download_all(self.handle, self._ptrs(outs, host_dtype), host_dtype),
              "d3h_batch_download_all")
        return outs

    def run(self, arith=N.ARITH_EXACT, depth=0, graph=True):
        t = N.Timing()
        check(lib().d3h_batch_run(self.handle, arith, depth, 1 if graph else 0, ctypes.byref(t)),
              "d3h_batch_run")
        return t.as_dict()

    def autotune(self, arith=N.ARITH_EXACT):
        """Choose depth, tile and z-chunks per size class for batched sweeps of
        this batch (the context keeps them).  Overwrites the fields: call it
        before uploading.  Returns the classes tuned and the 8-step whole-batch
        sweep time before and after."""
        n = ctypes.c_int()
        before, after = ctypes.c_double(), ctypes.c_double()
        check(lib().d3h_batch_autotune(self.handle, arith, ctypes.byref(n), ctypes.byref(before),
                                       ctypes.byref(after)), "d3h_batch_autotune")
        return {"classes": n.value, "sweep_ms_before": before.value, "sweep_ms_after": after.value}


def stable_coefficients(nx, ny, nz, kappa=0.1, dtype=np.float32):
    """Coefficients of the batched/served problems: dt = 0.1 h_min^2 / kappa with
    h = 1/n per axis, so every coefficient is <= 0.1 and cc >= 0.4 for any shape
    (the reference's dt uses dx only, diffusion3d/diffusion3d.h:61, which makes
    cc < 0 when dx^2/dy^2 + dx^2/dz^2 > 4).  Same float/double promotions as the
    reference constructor; for cubic grids the values equal the reference's."""
    if dtype == np.float64:
        dx, dy, dz = 1.0 / nx, 1.0 / ny, 1.0 / nz
        h = min(dx, dy, dz)
        dt = 0.1 * h * h / kappa
        ce, cn, ct = kappa * dt / (dx * dx), kappa * dt / (dy * dy), kappa * dt / (dz * dz)
        cc = 1.0 - (ce + ce + cn + cn + ct + ct)
        return [ce, ce, cn, cn, ct, ct, cc], dt
    f = np.float32
    kap = f(kappa)
    dx, dy, dz = f(1.0) / f(nx), f(1.0) / f(ny), f(1.0) / f(nz)
    h = min(dx, dy, dz)
    dt = f(0.1 * float(h) * float(h) / float(kap))
    ce = kap * dt / (dx * dx)
    cn = kap * dt / (dy * dy)
    ct = kap * dt / (dz * dz)
    s = f(f(f(f(f(ce + ce) + cn) + cn) + ct) + ct)
    cc = f(1.0 - float(s))
    return [float(v) for v in (ce, ce, cn, cn, ct, ct, cc)], float(dt)


class Server:
    """Continuous-batching request server (d3h_server_*)."""

    def __init__(self, ctx, hbm_budget_bytes=0, max_active=0, arith=N.ARITH_EXACT, depth=0,
                 sweep_steps=8):
        cfg = N.ServerConfig(hbm_budget_bytes, max_active, arith, depth, sweep_steps)
        h = ctypes.c_void_p()
        check(lib().d3h_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(h)),
              "d3h_server_create")
        self.ctx, self.handle = ctx, h
        self._keep = {}

    def close(self):
        if self.handle:
            lib().d3h_server_destroy(self.handle)
            self.handle = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, problem, host_in, host_dtype, out):
        """Queue a solve; `out` (numpy, dense) receives the result.  Both arrays
        are kept alive until the request is done."""
        p = N.Problem()
        p.nx, p.ny, p.nz = problem["nx"], problem["ny"], problem["nz"]
        p.dtype, p.steps = problem.get("dtype", N.F32), problem["steps"]
        for j, c in enumerate(problem["coef"]):
            p.coef[j] = float(c)
        p.rounding = problem.get("rounding", N.ROUND_NEAREST)
        p.seed = problem.get("seed", 0)
        out_dtype = {np.dtype(np.float32): N.F32, np.dtype(np.float64): N.F64,
                     np.dtype(np.float16): N.F16, np.dtype(np.uint16): N.BF16}[out.dtype]
        rid = ctypes.c_int64()
        check(lib().d3h_server_submit(self.handle, ctypes.byref(p), _dptr(host_in), host_dtype,
                                      _dptr(out), out_dtype, ctypes.byref(rid)), "d3h_server_submit")
        self._keep[rid.value] = (host_in, out)
        return rid.value

    def poll(self, rid):
        done, lat = ctypes.c_int(), ctypes.c_double()
        check(lib().d3h_server_poll(self.handle, rid, ctypes.byref(done), ctypes.byref(lat)),
              "d3h_server_poll")
        return bool(done.value), lat.value

    def wait(self, rid):
        lat = ctypes.c_double()
        check(lib().d3h_server_wait(self.handle, rid, ctypes.byref(lat)), "d3h_server_wait")
        self._keep.pop(rid, None)
        return lat.value

    def stats(self):
        c, r, s, b = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
        check(lib().d3h_server_stats(self.handle, ctypes.byref(c), ctypes.byref(r), ctypes.byref(s),
                                     ctypes.byref(b)), "d3h_server_stats")
        return {"completed": c.value, "rejected": r.value, "sweeps": s.value, "busy_ms": b.value}
