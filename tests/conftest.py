import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libvpt.so on cuda:0)")


@pytest.fixture(scope="session")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    from voxelpathtracer_amd import Context
    c = Context(0)
    yield c
    torch.cuda.synchronize()
    c.close()
