import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def gpu():
    """cuda:0 with the HIP library initialised; GPU tests fail loudly without it."""
    import torch

    import namsa_amd._lib as L

    assert torch.cuda.is_available(), "GPU test run without a visible HIP device"
    L.call("namsa_init", 0)
    return torch.device("cuda", 0)
