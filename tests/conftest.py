import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, 'gw-gpu-search_amd')
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP kernels through the C ABI)')


@pytest.fixture(scope='session')
def lib():
    import _lib
    if not os.path.exists(_lib.LIB_PATH):
        _lib.build()
    return _lib.load()


@pytest.fixture(scope='session')
def gpu(lib):
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a ROCm GPU'
    return torch.device('cuda')
