import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: long-running check")


@pytest.fixture(autouse=True)
def _torch_runtime_first(request):
    """The adjoint tests hand torch tensors to libathmi.so. torch brings its own
    HIP runtime, and it sees the GPU only if it initialises before the first
    HIP call through libathmi.so in the process, so every GPU test starts torch
    first, whatever order the tests run in."""
    if request.node.get_closest_marker("gpu") is not None:
        import torch
        torch.cuda.is_available()
    yield
