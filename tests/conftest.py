import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long CPU-only check")


def has_reference():
    return os.path.isdir(os.path.join(REFERENCE, "dim3"))


@pytest.fixture(scope="session")
def gpu():
    """Skip unless a GPU is visible; GPU tests must not silently fall back."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return 0
