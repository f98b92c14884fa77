import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmtgpu.so)")


@pytest.fixture(scope="session")
def dev():
    import torch
    import maroontree_amd as mt
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    mt.init(0)
    return torch.device("cuda:0")
