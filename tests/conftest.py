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
    nan_a, nan_b = np.isnan(a), np.isnan(b)
    if nan_a.any() or nan_b.any():
        if not np.array_equal(nan_a, nan_b):
            return np.inf
        m = ~nan_b
        if not m.any():
            return 0.0
        a, b = a[m], b[m]
    if b.size == 0:
        return 0.0
    ab = np.abs(b)
    scale = np.maximum(ab, floor_frac * ab.max() + 1e-300, out=ab)
    d = np.subtract(a, b)
    np.abs(d, out=d)
    np.divide(d, scale, out=d)
    return float(d.max())


def stencil_scale(v_rs, R):
    """Per face (ns, nk, nl, v): the largest |v| of its reconstruction stencil,
    J = j+3-R .. j+3+R of the reshaped input v_rs (ns+6, nk, nl, v).  A face
    value is a weighted sum of candidates of that size, so rounding in the sum
    is bounded relative to this scale, not to the face value itself (which can
    be small where candidates across a jump cancel)."""
    a = np.abs(np.asarray(v_rs, dtype=np.float64))
    ns = a.shape[0] - 6
    out = np.zeros((ns,) + a.shape[1:])
    for s in range(-R, R + 1):
        lo, hi = max(0, -(3 + s)), min(ns, a.shape[0] - (3 + s))
        out[lo:hi] = np.maximum(out[lo:hi], a[lo + 3 + s:hi + 3 + s])
    return out


def rel_err_stencil(a, b, scale):
    """max_i |a_i-b_i| / max(|b_i|, scale_i): rounding error measured against
    the size of the values the face was reconstructed from (stencil_scale)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        raise AssertionError(f"shape {a.shape} != {b.shape}")
    if not np.array_equal(np.isnan(a), np.isnan(b)):
        return np.inf
    m = ~np.isnan(b)
    return float(np.max(np.abs(a[m] - b[m]) / np.maximum(np.abs(b[m]), scale[m] + 1e-300)))


@pytest.fixture(scope="session")
def small_ref():
    return dict(np.load(os.path.join(GOLDEN, "ref_small.npz")))
