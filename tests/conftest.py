import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "oracle", "bitz-poc_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(autouse=True)
def _torch_hip_first(request):
    """GPU tests: torch's HIP state is initialised before the library's own
    HIP calls of the process (a test module that starts with the library's
    server threads would otherwise meet torch initialising afterwards)."""
    if request.node.get_closest_marker("gpu"):
        import torch
        torch.cuda.init()
    yield
