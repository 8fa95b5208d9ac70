import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dissolve-2.0_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built libdsv.so")


@pytest.fixture(scope="session")
def gpu():
    """Skips a GPU test when no GPU is visible; fails loudly if the library is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from dsv import _lib
    _lib.load()
    return torch.device("cuda", 0)
