import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")
    config.addinivalue_line("markers", "slow: large inputs")


@pytest.fixture(scope="session")
def dev():
    """The GPU device; GPU tests fail (not skip) when HIP or the library is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but no HIP device is visible")
    import gpumultisplit_amd as gms
    gms.load_library()
    return torch.device("cuda:0")
