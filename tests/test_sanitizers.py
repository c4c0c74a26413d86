"""Host-side sanitizers: the CPU oracle and the engine's host functions built
with gcc/g++ -fsanitize=address,undefined and driven over every function on
small odd-sized grids (no GPU).  Any report aborts the program and fails the
test."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HERE = os.path.dirname(os.path.abspath(__file__))
FLAGS = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-ffp-contract=off", "-fopenmp"]


def _build_and_run(tmp_path, compiler, sources, extra):
    if shutil.which(compiler) is None:
        pytest.skip("%s not available" % compiler)
    exe = str(tmp_path / "san")
    cmd = [compiler] + FLAGS + extra + sources + ["-lm", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", OMP_NUM_THREADS="2")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr


def test_oracle_under_asan_ubsan(tmp_path):
    _build_and_run(tmp_path, "gcc", [os.path.join(HERE, "native", "sanitize_oracle.c"),
                                     os.path.join(ROOT, "oracle", "oracle.c")],
                   ["-std=gnu11", "-I" + os.path.join(ROOT, "oracle")])


def test_host_functions_under_asan_ubsan(tmp_path):
    _build_and_run(tmp_path, "g++", [os.path.join(HERE, "native", "sanitize_host.cpp"),
                                     os.path.join(ROOT, "benchmarks_amd", "host", "initialize.cpp")],
                   ["-std=c++17", "-I" + os.path.join(ROOT, "include")])
