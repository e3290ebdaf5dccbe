import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950) and the built liboosg.so")


@pytest.fixture(scope="session")
def reference_lib():
    import _oracle
    lib = _oracle.reference()
    if lib is None:
        pytest.skip("the compiled reference (oracle/_ref) is not built here")
    return lib


@pytest.fixture(scope="session")
def gpu_device():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("a GPU test ran without a GPU")
    return torch.device("cuda:0")
