import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """The GPU-side wrappers; fails loudly if libfzh.so or the GPU is missing."""
    from fzgpumodules_amd import _fzh, kernels
    _fzh.lib()
    if _fzh.device_count() < 1:
        pytest.fail("no HIP device visible to libfzh.so")
    return kernels


def f32(x):
    return np.float32(x)


def eb_params(abs_eb):
    """(ebx2_r, ebx2) computed in float32 like lorenzo_quant.cu:643-644 / :511."""
    eb = np.float32(abs_eb)
    return np.float32(1) / (np.float32(2) * eb), np.float32(2) * eb
