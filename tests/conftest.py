import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libfv3atm_amd.so")


@pytest.fixture(scope="session")
def gpu_ready():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a host without a HIP device")
    from fv3atm_amd import _ffi
    _ffi.lib()  # fails loudly if the extension is missing
    return True
