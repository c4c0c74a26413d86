"""ctypes declarations of the engine's C ABI (include/diffusion3d_hip.h).

The shared library is built in-tree (benchmarks_amd/lib/libdiffusion3d_hip.so)
by ``make -C benchmarks_amd/csrc``.  There is no fallback: if the library is
missing, :func:`lib` raises.
"""
import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("D3H_LIB") or os.path.join(HERE, "lib", "libdiffusion3d_hip.so")
DRIVER_PATH = os.path.join(HERE, "bin", "diffusion3d_hip")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "diffusion3d_hip.h")

D3H_OK = 0
D3H_ERR_INVALID = 1
D3H_ERR_HIP = 2
D3H_ERR_NOMEM = 3
D3H_ERR_UNSUPPORTED = 4
D3H_ERR_STATE = 5
D3H_ERR_NOT_FOUND = 6
D3H_ERR_UNSTABLE = 7

F32, F64, BF16, F16 = 0, 1, 2, 3
ARITH_EXACT, ARITH_FMA = 0, 1
ROUND_NEAREST, ROUND_STOCHASTIC = 0, 1

c_int, c_double, c_float, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_void_p
c_size_t, c_int64, c_uint64, c_char_p = ctypes.c_size_t, ctypes.c_int64, ctypes.c_uint64, ctypes.c_char_p
c_uint = ctypes.c_uint
P = ctypes.POINTER


class RunConfig(ctypes.Structure):
    _fields_ = [("temporal_depth", c_int), ("arith", c_int), ("tile", c_int),
                ("zchunks", c_int), ("use_graph", c_int), ("rounding", c_int), ("seed", c_uint),
                ("pace_ns", c_int)]


class Timing(ctypes.Structure):
    _fields_ = [("kernel_ms", c_double), ("wall_ms", c_double), ("avg_launch_ms", c_double),
                ("launches", c_int), ("steps", c_int), ("temporal_depth", c_int),
                ("tile", c_int), ("zchunks", c_int), ("blocks", c_int), ("pace_ns", c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TileDesc(ctypes.Structure):
    _fields_ = [("dtype", c_int), ("temporal_depth", c_int), ("arith", c_int), ("groups", c_int),
                ("rows_per_wave", c_int), ("waves", c_int), ("tile_x", c_int), ("tile_y", c_int),
                ("prefetch", c_int), ("row_stream", c_int)]


class Problem(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("nz", c_int), ("dtype", c_int), ("steps", c_int),
                ("coef", c_double * 7), ("rounding", c_int), ("seed", c_uint)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", c_size_t), ("max_active", c_int), ("arith", c_int),
                ("temporal_depth", c_int), ("sweep_steps", c_int)]


# name -> (restype, argtypes)
SIGNATURES = {
    "initialize_hip": (c_int, [c_int, P(c_char_p), c_int, c_int, c_int]),
    "initialize_benchmark_hip": (c_int, [c_int, c_int, c_int]),
    "run_kernel_hip": (c_int, [c_int, P(c_float), c_int, c_int, c_int] + [c_float] * 7),
    "finalize_benchmark_hip": (c_int, []),
    "kernel_time_hip": (c_int, [P(c_float)]),
    "diffusion_hip_": (c_int, [P(c_float), P(c_float), P(c_int), P(c_int), P(c_int)] + [P(c_float)] * 7
                       + [P(c_int)]),
    "d3h_status_string": (c_char_p, [c_int]),
    "d3h_last_error": (c_char_p, []),
    "d3h_context_create": (c_int, [c_int, P(c_void_p)]),
    "d3h_context_destroy": (c_int, [c_void_p]),
    "d3h_device_info": (c_int, [c_void_p, P(c_int), P(c_size_t), P(c_size_t), P(c_int)]),
    "d3h_host_alloc": (c_int, [c_size_t, P(c_void_p)]),
    "d3h_host_free": (c_int, [c_void_p]),
    "d3h_run_config_default": (None, [P(RunConfig)]),
    "d3h_field_create": (c_int, [c_void_p, c_int, c_int, c_int, c_int, P(c_void_p)]),
    "d3h_field_destroy": (c_int, [c_void_p]),
    "d3h_field_dims": (c_int, [c_void_p, P(c_int), P(c_int), P(c_int), P(c_int)]),
    "d3h_field_upload": (c_int, [c_void_p, c_void_p, c_int]),
    "d3h_field_download": (c_int, [c_void_p, c_void_p, c_int]),
    "d3h_field_copy": (c_int, [c_void_p, c_void_p]),
    "d3h_field_save": (c_int, [c_void_p, c_char_p]),
    "d3h_field_load": (c_int, [c_void_p, c_char_p]),
    "d3h_raw_write": (c_int, [c_char_p, c_void_p, c_int, c_int, c_int, c_int, c_uint64, c_int]),
    "d3h_raw_read": (c_int, [c_char_p, c_void_p, c_size_t, P(c_int), P(c_int), P(c_int), P(c_int),
                             P(c_uint64), P(c_int)]),
    "d3h_field_set_y_reversed": (c_int, [c_void_p, c_int]),
    "d3h_field_run": (c_int, [c_void_p, c_int, P(c_double), P(RunConfig), P(Timing)]),
    "d3h_field_compare": (c_int, [c_void_p, c_void_p, P(c_double)]),
    "d3h_field_error_analytical": (c_int, [c_void_p, c_double, c_double, P(c_double), P(c_double),
                                           P(c_double)]),
    "d3h_field_error_analytical_noalloc": (c_int, [c_void_p, c_double, c_double, P(c_double), P(c_double),
                                           P(c_double)]),
    "d3h_tile_count": (c_int, []),
    "d3h_tile_info": (c_int, [c_int, P(TileDesc)]),
    "d3h_autotune": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, P(c_int), P(c_int),
                             P(c_double)]),
    "d3h_autotune_save": (c_int, [c_void_p, c_char_p]),
    "d3h_autotune_load": (c_int, [c_void_p, c_char_p]),
    "d3h_bandwidth_probe": (c_int, [c_void_p, c_size_t, c_int, c_int, P(c_double), P(c_double)]),
    "d3h_constants_f32": (c_int, [c_int, c_int, c_int, P(c_float)]),
    "d3h_constants_f64": (c_int, [c_int, c_int, c_int, P(c_double)]),
    "d3h_initialize_f32": (c_int, [P(c_float), c_int, c_int, c_int, P(c_float), c_float]),
    "d3h_initialize_f64": (c_int, [P(c_double), c_int, c_int, c_int, P(c_double), c_double]),
    "d3h_constants2d_f32": (c_int, [c_int, c_int, P(c_float)]),
    "d3h_initialize2d_f32": (c_int, [P(c_float), c_int, c_int, P(c_float), c_float]),
    "d3h_add_noise_f32": (c_int, [P(c_float), c_size_t, c_uint64, c_float]),
    "d3h_problem_check": (c_int, [P(Problem)]),
    "d3h_problem_footprint": (c_int, [P(Problem), c_int, c_int, P(c_size_t)]),
    "d3h_batch_create": (c_int, [c_void_p, P(Problem), c_int, P(c_void_p)]),
    "d3h_batch_destroy": (c_int, [c_void_p]),
    "d3h_batch_upload": (c_int, [c_void_p, c_int, c_void_p, c_int]),
    "d3h_batch_download": (c_int, [c_void_p, c_int, c_void_p, c_int]),
    "d3h_batch_upload_all": (c_int, [c_void_p, P(c_void_p), c_int]),
    "d3h_batch_download_all": (c_int, [c_void_p, P(c_void_p), c_int]),
    "d3h_batch_run": (c_int, [c_void_p, c_int, c_int, c_int, P(Timing)]),
    "d3h_batch_autotune": (c_int, [c_void_p, c_int, P(c_int), P(c_double), P(c_double)]),
    "d3h_server_create": (c_int, [c_void_p, P(ServerConfig), P(c_void_p)]),
    "d3h_server_destroy": (c_int, [c_void_p]),
    "d3h_server_submit": (c_int, [c_void_p, P(Problem), c_void_p, c_int, c_void_p, c_int,
                                  P(c_int64)]),
    "d3h_server_poll": (c_int, [c_void_p, c_int64, P(c_int), P(c_double)]),
    "d3h_server_wait": (c_int, [c_void_p, c_int64, P(c_double)]),
    "d3h_server_stats": (c_int, [c_void_p, P(c_int64), P(c_int64), P(c_int64), P(c_double)]),
}

_lock = threading.Lock()
_lib = None


class D3HError(RuntimeError):
    def __init__(self, status, where, message):
        super().__init__("%s: %s (%s)" % (where, message, STATUS_NAMES.get(status, status)))
        self.status = status


STATUS_NAMES = {0: "OK", 1: "INVALID", 2: "HIP", 3: "NOMEM", 4: "UNSUPPORTED", 5: "STATE",
                6: "NOT_FOUND", 7: "UNSTABLE"}


def lib():
    """The loaded engine library (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError("HIP engine library not built: %s (run `make -C "
                                   "benchmarks_amd/csrc`)" % LIB_PATH)
            handle = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def check(status, where):
    if status != D3H_OK:
        msg = lib().d3h_last_error().decode(errors="replace")
        raise D3HError(status, where, msg)
    return status


def header_functions(path=HEADER_PATH):
    """Names of the functions declared in include/diffusion3d_hip.h."""
    import re
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*([a-z_][a-z0-9_]*)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))
