import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def ctx():
    # torch carries its own HIP runtime; it has to initialise the device before
    # the library's runtime does, or torch finds no GPU in this process
    import torch

    torch.cuda.init()
    from its_amd import _capi

    return _capi.Context(0)
