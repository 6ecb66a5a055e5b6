import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "cans-fizzy_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running physics check")


def halo_shape(n):
    return (n[0] + 2, n[1] + 2, n[2] + 2)


def rand_halo(rng, n, scale=1.0, offset=0.0):
    return np.asfortranarray(offset + scale * rng.standard_normal(halo_shape(n)))


def rand_interior(rng, n, scale=1.0):
    return np.asfortranarray(scale * rng.standard_normal(tuple(n)))


def rel_err(a, b):
    """max|a-b| / max|b| (0 when both vanish)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.abs(b).max()
    num = np.abs(a - b).max()
    if den == 0.0:
        return float(num)
    return float(num / den)


def demean(a):
    a = np.asarray(a, dtype=np.float64)
    return a - a.mean()


@pytest.fixture
def rng():
    return np.random.default_rng(20261202)
