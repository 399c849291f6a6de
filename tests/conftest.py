import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def corc():
    """The C FP32 restatement (built on first use)."""
    from oracle import c_oracle
    c_oracle.lib()
    return c_oracle


@pytest.fixture(scope="session")
def nporc():
    from oracle import np_oracle
    return np_oracle


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import pde_module_amd._native as N
    N.load()
    return torch.device("cuda:0")


def same_bits(a, b, mask=None):
    """Bitwise equality of two FP32 arrays (NaN == NaN), optionally on a mask."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    if mask is not None:
        a, b = a[mask], b[mask]
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))
