"""Write golden fixtures from the reference's own CPU build (oracle/_ref/libref.so).

Run in the build container only (the reference tree is mounted there):

    make -C oracle ref && python tests/golden/gen_golden.py

The library is compiled by oracle/Makefile from the reference sources where
they lie (src/diffusion3d/{diffusion3d,baseline,diffusion3d_openmp,
diffusion3d_openmp_temporal_blocking}.cc) plus oracle/ref_probe.cc.  The
fixtures are data only: raw float32 outputs of Baseline::RunKernel,
Diffusion3DOpenMP::RunKernel and Diffusion3DOpenMPTemporalBlocking::RunKernel,
the constructor's constants, GetCorrectAnswer fields and GetAccuracy values.
Large outputs are recorded as SHA-256 of the raw little-endian bytes plus
strided samples.
"""
import ctypes
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref.so")

FULL_SHAPES = [(8, 8, 8), (16, 16, 16), (12, 10, 7), (33, 17, 5), (5, 9, 31)]
HASH_SHAPES = [(32, 32, 32), (128, 128, 128)]
COUNTS = [0, 1, 2, 7, 100]
CONST_SHAPES = [(8, 8, 8), (16, 16, 16), (32, 32, 32), (128, 128, 128), (256, 256, 256),
                (512, 512, 512), (768, 768, 768), (12, 10, 7), (33, 17, 5), (5, 9, 31),
                (64, 32, 16)]
SAMPLE_STRIDE = 997


def _lib():
    lib = ctypes.CDLL(LIB)
    fp = ctypes.POINTER(ctypes.c_float)
    lib.ref_run.argtypes = [ctypes.c_int] * 5 + [fp, fp, fp, fp, ctypes.POINTER(ctypes.c_double)]
    lib.ref_run.restype = ctypes.c_int
    lib.ref_analytical.argtypes = [ctypes.c_int] * 4 + [fp]
    lib.ref_analytical.restype = ctypes.c_int
    lib.ref_constants.argtypes = [ctypes.c_int] * 3 + [fp]
    lib.ref_constants.restype = ctypes.c_int
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def run(lib, variant, shape, count):
    nx, ny, nz = shape
    out = np.empty(nz * ny * nx, dtype=np.float32)
    acc = np.zeros(1, dtype=np.float32)
    const = np.zeros(15, dtype=np.float32)
    sec = ctypes.c_double(0)
    rc = lib.ref_run(variant, nx, ny, nz, count, None, _ptr(out), _ptr(acc), _ptr(const),
                     ctypes.byref(sec))
    assert rc == 0
    return out, float(acc[0]), const


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    lib = _lib()
    gomp = ctypes.CDLL("libgomp.so.1")
    gomp.omp_set_num_threads(8)
    arrays = {}
    meta = {"source": "reference CPU build (oracle/_ref/libref.so)", "cases": [],
            "constants": {}, "tb": [], "analytical": []}

    for shape in CONST_SHAPES:
        nx, ny, nz = shape
        const = np.zeros(15, dtype=np.float32)
        assert lib.ref_constants(nx, ny, nz, _ptr(const)) == 0
        meta["constants"]["x".join(map(str, shape))] = [int(v) for v in const.view(np.uint32)]

    for shape in FULL_SHAPES + HASH_SHAPES:
        for count in COUNTS:
            if shape == (128, 128, 128) and count not in (0, 1, 100):
                continue
            serial, acc_s, _ = run(lib, 0, shape, count)
            omp, acc_o, _ = run(lib, 1, shape, count)
            assert np.array_equal(serial.view(np.uint32), omp.view(np.uint32)), (shape, count)
            key = "x".join(map(str, shape)) + "_c%d" % count
            case = {"shape": list(shape), "count": count, "sha256": sha(serial),
                    "accuracy_bits": int(np.float32(acc_s).view(np.uint32)),
                    "accuracy": "%e" % acc_s}
            if shape in FULL_SHAPES:
                arrays["out_" + key] = serial
            else:
                idx = np.arange(0, serial.size, SAMPLE_STRIDE)
                case["sample_stride"] = SAMPLE_STRIDE
                case["samples_bits"] = [int(v) for v in serial[idx].view(np.uint32)]
            meta["cases"].append(case)
            print("case", key, case["accuracy"], flush=True)

    # Known answer printed by the reference at its defaults (256^3, 100 steps).
    out, acc, _ = run(lib, 1, (256, 256, 256), 100)
    meta["default_256_c100"] = {"sha256": sha(out), "accuracy": "%e" % acc,
                                "accuracy_bits": int(np.float32(acc).view(np.uint32))}
    print("256^3/100 accuracy", "%e" % acc, flush=True)

    # OpenMP temporal blocking: bitwise equal to plain for even counts.
    for threads in (2, 4, 8):
        gomp.omp_set_num_threads(threads)
        for count in (2, 100):
            tb, _, _ = run(lib, 2, (32, 32, 32), count)
            meta["tb"].append({"threads": threads, "count": count, "shape": [32, 32, 32],
                               "sha256": sha(tb)})
    gomp.omp_set_num_threads(8)

    # Analytical fields at t = count*dt (GetCorrectAnswer).
    for shape in [(8, 8, 8), (12, 10, 7), (33, 17, 5)]:
        for count in (0, 7, 100, 1000):
            nx, ny, nz = shape
            a = np.empty(nx * ny * nz, dtype=np.float32)
            assert lib.ref_analytical(nx, ny, nz, count, _ptr(a)) == 0
            key = "ana_" + "x".join(map(str, shape)) + "_c%d" % count
            arrays[key] = a
            meta["analytical"].append({"shape": list(shape), "count": count, "key": key})

    np.savez_compressed(os.path.join(HERE, "reference_cpu.npz"), **arrays)
    with open(os.path.join(HERE, "reference_cpu.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print("wrote", len(arrays), "arrays")


if __name__ == "__main__":
    sys.exit(main())
