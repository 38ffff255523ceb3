"""Shared test set-up: import paths, the ``gpu`` marker, fixture loading."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "softx-d-17-00015_amd")
for p in (os.path.join(REPO, "oracle"), PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def load_golden(name):
    path = os.path.join(GOLDEN, f"golden_{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"fixture {path} not present")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel_err(a, b):
    """Normwise relative error max|a-b| / max|b| (per field)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))
