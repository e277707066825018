import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X (gfx950)")
    config.addinivalue_line("markers", "slow: CPU oracle runs that take tens of seconds")


def _have_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """cuda:0 as a torch device (the GPU tests call the library through the C ABI;
    torch only allocates device memory)."""
    if not _have_gpu():
        pytest.skip("no HIP device")
    import torch
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def pg():
    import gpuintegration_amd
    gpuintegration_amd.load()
    return gpuintegration_amd
