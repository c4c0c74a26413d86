"""The reference-side binding: the HIP engine as one more variant of the
reference's diffusion3d benchmark (tests/native/reference_variant, the class a
maintainer adds as src/diffusion3d/diffusion3d_hip.h, following
diffusion3d/diffusion3d_physis.h:21-43, and the main.cc branch beside
diffusion3d/main.cc:163).

CPU: the variant compiles against the reference's own headers and its
diffusion3d.cc / baseline.cc with the reference's flags, links to the engine
library, and every engine symbol it needs is exported by the library (skipped
where the reference tree is not mounted).

GPU: the variant built that way (oracle/_ref/diffusion3d_hip_variant.exe) and
the repository's own driver print the reference OpenMP build's Accuracy line
and write its "%.7E" dump (diffusion3d/baseline.cc:62-70) byte for byte; the
driver's dump is also checked line by line against the oracle.
"""
import os
import re
import subprocess

import pytest

from conftest import ROOT

REF_SRC = "/root/reference/src"
VARIANT = os.path.join(ROOT, "tests", "native", "reference_variant")
LIBDIR = os.path.join(ROOT, "benchmarks_amd", "lib")
REF_BIN = os.path.join(ROOT, "oracle", "_ref")
# the reference's flags (common/common.mk:17, common/gcc.mk:22)
REF_FLAGS = ["--std=c++11", "-O3", "-Wall"]


def _exported(lib_path):
    r = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True,
                       check=True)
    return {l.split()[-1] for l in r.stdout.splitlines() if l.strip()}


def test_variant_compiles_and_links_against_reference(tmp_path):
    if not os.path.isdir(os.path.join(REF_SRC, "diffusion3d")):
        pytest.skip("reference tree not mounted")
    lib = os.path.join(LIBDIR, "libdiffusion3d_hip.so")
    if not os.path.exists(lib):
        pytest.skip("engine library not built")
    exe = str(tmp_path / "diffusion3d_hip_variant.exe")
    cmd = (["g++"] + REF_FLAGS + ["-I" + REF_SRC, "-I" + VARIANT, "-I" + os.path.join(ROOT, "include"),
            "-o", exe, os.path.join(VARIANT, "hip_main.cc"),
            os.path.join(REF_SRC, "diffusion3d", "diffusion3d.cc"),
            os.path.join(REF_SRC, "diffusion3d", "baseline.cc"), "-lm", "-L" + LIBDIR,
            "-ldiffusion3d_hip", "-Wl,-rpath," + LIBDIR])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "warning" not in r.stderr.lower(), r.stderr
    undef = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True,
                           check=True).stdout
    needed = {l.split()[-1] for l in undef.splitlines()
              if re.search(r"\b(\w+_hip|d3h_\w+)$", l.strip())}
    assert {"initialize_hip", "initialize_benchmark_hip", "run_kernel_hip",
            "finalize_benchmark_hip", "kernel_time_hip", "d3h_status_string",
            "d3h_last_error"} <= needed, needed
    assert needed <= _exported(lib), needed - _exported(lib)


def _run(exe, args, cwd):
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=300, cwd=cwd)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _accuracy(out):
    return re.search(r"Accuracy\s+:\s+(\S+)", out).group(1)


@pytest.mark.gpu
def test_driver_dump_matches_oracle(tmp_path):
    """diffusion3d_hip --size 64 --count 20 --dump: every "%.7E" line of
    diffusion3d_result.hip.out is the oracle's value after 20 steps."""
    import benchmarks_amd as d3
    import oracle
    from benchmarks_amd import _native as N
    out = _run(N.DRIVER_PATH, ["--size", "64", "--count", "20", "--dump"], tmp_path)
    k = d3.constants_f32(64, 64, 64)
    want = oracle.run_f32(d3.initialize_f32(64, 64, 64, k), k[8:15], 20)
    assert _accuracy(out) == "%e" % oracle.accuracy_f32(want, k, 20), out
    lines = (tmp_path / "diffusion3d_result.hip.out").read_text().splitlines()
    assert len(lines) == 64 ** 3
    assert lines == ["%.7E" % float(v) for v in want.ravel()]


@pytest.mark.gpu
@pytest.mark.parametrize("size,count", [(64, 20), (48, 7)])
def test_variant_and_driver_equal_reference_build(tmp_path, size, count):
    """The variant linked into the reference's own classes, and the
    repository's driver, against the reference's OpenMP build: same Accuracy
    line, byte-identical dump files."""
    ref_exe = os.path.join(REF_BIN, "diffusion3d_openmp.exe")
    var_exe = os.path.join(REF_BIN, "diffusion3d_hip_variant.exe")
    if not (os.path.exists(ref_exe) and os.path.exists(var_exe)):
        pytest.skip("reference build (oracle/_ref, `make -C oracle ref`) not present")
    from benchmarks_amd import _native as N
    args = ["--size", str(size), "--count", str(count), "--dump"]
    dirs = {}
    outs = {}
    for name, exe in (("ref", ref_exe), ("variant", var_exe), ("driver", N.DRIVER_PATH)):
        dirs[name] = tmp_path / name
        dirs[name].mkdir()
        outs[name] = _run(exe, args, dirs[name])
    acc = _accuracy(outs["ref"])
    assert _accuracy(outs["variant"]) == acc and _accuracy(outs["driver"]) == acc, outs
    assert "Running diffusion3d/hip" in outs["variant"] and "Kernel-only performance:" in outs["variant"]
    ref_dump = (dirs["ref"] / "diffusion3d_result.openmp.out").read_bytes()
    assert len(ref_dump.splitlines()) == size ** 3
    assert (dirs["variant"] / "diffusion3d_result.hip.out").read_bytes() == ref_dump
    assert (dirs["driver"] / "diffusion3d_result.hip.out").read_bytes() == ref_dump
