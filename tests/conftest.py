import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libwhhd.so")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def whhd():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test without a HIP device")
    import waveletpotentials_amd as W
    W.lib()  # raises if libwhhd.so is missing: there is no fallback path
    return W


@pytest.fixture(params=[0, 1], ids=["line_parallel", "windowed"])
def kernel_mode(request, whhd):
    """Run a GPU test with the line-parallel kernels enabled (where they exist)
    and again with the windowed tile kernels only."""
    L = whhd.lib()
    assert L.whhd_set_line_kernel(request.param) == 0
    yield request.param
    L.whhd_set_line_kernel(0)
