import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libvesper_hip kernels)")


@pytest.fixture(scope="session")
def gpu():
    """The GPU device; fails (never skips) when no GPU or library is present,
    so a GPU run cannot pass on a silent fallback."""
    import torch
    from vesper_python_amd import _native
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    _native.load()
    return torch.device("cuda", 0)


@pytest.fixture(scope="session")
def fastvec_ref():
    """ctypes handle of the oracle's C restatement of doFastVEC."""
    from oracle import build_c
    return build_c.load()
