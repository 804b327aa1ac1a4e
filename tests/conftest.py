import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs a HIP device (MI355X); runs the HIP library through the C-ABI")


@pytest.fixture(scope="session")
def torch_gpu():
    import torch

    assert torch.cuda.is_available(), "GPU test run without a visible HIP device"
    return torch
