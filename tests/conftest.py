import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU guard: a gpu-marked test must run where a GPU is visible."""
    import torch
    assert torch.cuda.is_available(), "gpu tests need a visible ROCm device"
    from latticeboltzmann_amd import _native
    _native.load()
    return torch.device("cuda", 0)
