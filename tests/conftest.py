import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def rel_err(a, b, floor_frac=1e-3):
    """max_i |a_i-b_i| / max(|b_i|, floor_frac*max|b|): element-wise relative
    error with a floor that keeps values near zero from dividing by ~0."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        raise AssertionError(f"shape {a.shape} != {b.shape}")
    if not np.array_equal(np.isnan(a), np.isnan(b)):
        return np.inf
    m = ~np.isnan(b)
    if not m.any():
        return 0.0
    scale = np.maximum(np.abs(b[m]), floor_frac * np.max(np.abs(b[m])) + 1e-300)
    return float(np.max(np.abs(a[m] - b[m]) / scale))


@pytest.fixture(scope="session")
def small_ref():
    return dict(np.load(os.path.join(GOLDEN, "ref_small.npz")))
