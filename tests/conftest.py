import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libmxllg.so")


def golden(name):
    import numpy as np
    return np.load(os.path.join(GOLDEN, name))


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU guard: GPU tests must run the HIP path, never fall back."""
    import torch
    assert torch.cuda.is_available(), "GPU test collected on a machine without a HIP device"
    from magex_amd import _lib
    _lib.lib()
    return torch
