"""MI355X-native 3D 7-point diffusion engine.

Hand-written HIP kernels for gfx950 behind a C ABI (include/diffusion3d_hip.h),
a C++ benchmark-class driver (benchmarks_amd/host) and this Python binding.
"""
from ._native import (ARITH_EXACT, ARITH_FMA, BF16, F16, F32, F64, ROUND_NEAREST, ROUND_STOCHASTIC,
                      D3HError, LIB_PATH, lib)
from .engine import (Batch, Context, Field, Server, stable_coefficients, add_noise_f32,
                     coefficients, constants_f32, constants_f64, constants2d_f32,
                     initialize_f32, initialize_f64, initialize2d_f32, load_raw, run_kernel,
                     run_kernel2d, save_raw, tiles)

__all__ = ["ARITH_EXACT", "ARITH_FMA", "BF16", "F16", "F32", "F64", "ROUND_NEAREST",
           "ROUND_STOCHASTIC", "D3HError", "LIB_PATH", "lib",
           "Batch", "Context", "Field", "Server", "stable_coefficients", "add_noise_f32",
           "coefficients", "constants_f32", "constants_f64", "constants2d_f32",
           "initialize_f32", "initialize_f64", "initialize2d_f32", "load_raw", "run_kernel",
           "run_kernel2d", "save_raw", "tiles"]
