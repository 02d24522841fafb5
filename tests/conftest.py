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
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libfdtd_hip.so")


def same_bits(a, b) -> bool:
    """True when a and b hold the same values bit for bit: floating-point
    arrays of one dtype are compared as their integer bit patterns (so +0 and
    -0 differ, and a NaN equals the same NaN); anything else by value."""
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype == b.dtype and a.dtype.kind == "f":
        u = np.dtype(f"u{a.dtype.itemsize}")
        return bool(np.array_equal(np.ascontiguousarray(a).view(u), np.ascontiguousarray(b).view(u)))
    return bool(np.array_equal(a, b))


def golden(name: str):
    """(meta dict, {key: array}) of tests/golden/<name>.npz."""
    d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    meta = json.loads(str(d["meta"]))
    return meta, {k: d[k] for k in d.files if k != "meta"}


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    from simulation_amd import _lib
    _lib.lib()
    return torch.device("cuda", 0)
