import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "matmul_golden.json")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def real_mat(rows):
    """Parse a transcribed real literal (list of rows or flat list) with float(),
    which rounds correctly like Java's Double.parseDouble."""
    if rows and isinstance(rows[0], list):
        return np.array([[float(x) for x in r] for r in rows], dtype=np.float64)
    return np.array([float(x) for x in rows], dtype=np.float64)


def cplx_mat(rows):
    if rows and isinstance(rows[0][0], list):
        return np.array([[complex(float(re), float(im)) for re, im in r] for r in rows], dtype=np.complex128)
    return np.array([complex(float(re), float(im)) for re, im in rows], dtype=np.complex128)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)
