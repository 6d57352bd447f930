import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")
    # the oracle's linear algebra is many small matrices: one BLAS thread per process is the fastest, and
    # several test processes with a full BLAS thread pool each oversubscribe the cores by far
    try:
        from threadpoolctl import threadpool_limits
        threadpool_limits(1, user_api="blas")
    except ImportError:
        pass


def rel_err(a, ref):
    """max |a - ref| / max |ref| over the whole array (the error measure used throughout)."""
    a = np.asarray(a)
    ref = np.asarray(ref)
    den = np.abs(ref).max()
    return float(np.abs(a - ref).max() / (den if den > 0 else 1.0))


def load_golden(name):
    path = os.path.join(GOLDEN, name + ".npz")
    return dict(np.load(path, allow_pickle=False))


@pytest.fixture(scope="session")
def gpu_available():
    import torch
    return torch.cuda.is_available()
