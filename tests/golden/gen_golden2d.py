"""Write golden fixtures for the diffusion/ family (2D 5-point operator and its
3D form with y-1 summed before y+1) from the reference's own CPU build
(oracle/_ref/libref2d.so).

Run in the build container only (the reference tree is mounted there):

    make -C oracle ref && python tests/golden/gen_golden2d.py

The library is compiled by oracle/Makefile from src/diffusion/diffusion_openmp.cc
and the family's headers (diffusion.h, baseline.h) where they lie, plus
oracle/ref_probe2d.cc.  The fixtures are data only: raw float32 outputs of
Baseline::RunKernel and DiffusionOpenMP::RunKernel, the constructor's
constants and GetAccuracy values.  Large outputs are recorded as SHA-256 of
the raw little-endian bytes plus strided samples.
"""
import ctypes
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref2d.so")

FULL_2D = [(8, 8), (16, 16), (13, 7), (33, 17), (5, 31), (64, 64), (1, 9), (9, 1)]
HASH_2D = [(256, 256), (1024, 1024)]
COUNTS = [0, 1, 2, 7, 100]
CONST_2D = [(8, 8), (16, 16), (13, 7), (33, 17), (5, 31), (64, 64), (256, 256),
            (1024, 1024), (4096, 4096), (8192, 8192)]
FULL_3D = [(8, 8, 8), (12, 10, 7), (33, 17, 5)]
HASH_3D = [(64, 64, 64)]
SAMPLE_STRIDE = 997


def _lib():
    lib = ctypes.CDLL(LIB)
    fp = ctypes.POINTER(ctypes.c_float)
    lib.ref2d_run.argtypes = [ctypes.c_int] * 6 + [fp, fp, fp, fp, ctypes.POINTER(ctypes.c_double)]
    lib.ref2d_run.restype = ctypes.c_int
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def run(lib, variant, nd, shape, count):
    nx, ny, nz = (tuple(shape) + (1,))[:3]
    out = np.empty(nz * ny * nx, dtype=np.float32)
    acc = np.zeros(1, dtype=np.float32)
    const = np.zeros(15, dtype=np.float32)
    rc = lib.ref2d_run(variant, nd, nx, ny, nz, count, None, _ptr(out), _ptr(acc), _ptr(const), None)
    assert rc == 0
    return out, float(acc[0]), const


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def bits(a):
    return [int(v) for v in np.asarray(a, np.float32).view(np.uint32)]


def main():
    lib = _lib()
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(8)
    arrays = {}
    meta = {"source": "reference CPU build (oracle/_ref/libref2d.so), src/diffusion",
            "constants2d": {}, "cases2d": [], "cases3d": []}

    for shape in CONST_2D:
        _, _, const = run(lib, 1, 2, shape, 0)
        meta["constants2d"]["x".join(map(str, shape))] = bits(const)

    for shape in FULL_2D + HASH_2D:
        for count in COUNTS:
            serial, acc_s, const = run(lib, 0, 2, shape, count)
            omp, acc_o, _ = run(lib, 1, 2, shape, count)
            assert np.array_equal(serial.view(np.uint32), omp.view(np.uint32)), (shape, count)
            key = "x".join(map(str, shape)) + "_c%d" % count
            case = {"shape": list(shape), "count": count, "sha256": sha(serial),
                    "accuracy_bits": bits([acc_s])[0], "accuracy": "%e" % acc_s,
                    "constants_bits": bits(const)}
            if shape in FULL_2D:
                arrays["out2d_" + key] = serial
            else:
                idx = np.arange(0, serial.size, SAMPLE_STRIDE)
                case["sample_stride"] = SAMPLE_STRIDE
                case["samples_bits"] = bits(serial[idx])
            meta["cases2d"].append(case)
            print("2d", key, case["accuracy"], flush=True)

    # The family's 3D form (DiffusionOpenMP, ndim = 3): y-1 before y+1, and
    # cc = 1 - (ce + cw + cn + cs + (ct + cb)).
    for shape in FULL_3D + HASH_3D:
        for count in (0, 1, 7, 100):
            omp, acc_o, const = run(lib, 1, 3, shape, count)
            key = "x".join(map(str, shape)) + "_c%d" % count
            case = {"shape": list(shape), "count": count, "sha256": sha(omp),
                    "accuracy_bits": bits([acc_o])[0], "accuracy": "%e" % acc_o,
                    "constants_bits": bits(const)}
            if shape in FULL_3D:
                arrays["out3d_" + key] = omp
            else:
                idx = np.arange(0, omp.size, SAMPLE_STRIDE)
                case["sample_stride"] = SAMPLE_STRIDE
                case["samples_bits"] = bits(omp[idx])
            meta["cases3d"].append(case)
            print("3d", key, case["accuracy"], flush=True)

    np.savez_compressed(os.path.join(HERE, "reference_cpu_2d.npz"), **arrays)
    with open(os.path.join(HERE, "reference_cpu_2d.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print("wrote", len(arrays), "arrays")


if __name__ == "__main__":
    sys.exit(main())
