import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def lib():
    from mrhs_qcu_amd import _lib

    return _lib.load()


@pytest.fixture(scope="session")
def dev():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("GPU test without a visible GPU")
    return torch.device("cuda:0")
