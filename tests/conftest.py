import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def golden():
    """Fixtures written from the reference's own CPU build (tests/golden/gen_golden.py)."""
    with open(os.path.join(GOLDEN, "reference_cpu.json")) as f:
        meta = json.load(f)
    arrays = dict(np.load(os.path.join(GOLDEN, "reference_cpu.npz")))
    return meta, arrays


@pytest.fixture(scope="session")
def ctx():
    import benchmarks_amd as d3
    c = d3.Context(0)
    yield c
    c.close()


def shape_key(shape):
    return "x".join(str(v) for v in shape)


def as_grid(flat, shape):
    nx, ny, nz = shape
    return np.asarray(flat).reshape(nz, ny, nx)
