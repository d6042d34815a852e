import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tarang.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    return load


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip-free on a machine with a GPU: a gpu-marked test that finds no device fails."""
    import torch
    assert torch.cuda.is_available(), "gpu test run without a visible HIP device"
    from tests import gpu_helpers
    return gpu_helpers.Device()


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)


def staterr(pairs):
    """Largest error over several fields of one state, relative to the state's largest
    entry: the measure for fields far smaller than the others (a velocity 1e-5 of the
    buoyancy), whose own-scale error is set by the rounding of the large fields."""
    num = max(float(np.abs(np.asarray(a) - np.asarray(b)).max()) for a, b in pairs)
    den = max(float(np.abs(np.asarray(b)).max()) for _, b in pairs)
    return num / max(den, 1e-300)
