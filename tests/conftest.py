import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libnbx.so on cuda:0)")


@pytest.fixture(scope="session")
def nbo():
    import nbo as _nbo
    _nbo.lib()
    return _nbo


@pytest.fixture(scope="session")
def nbx():
    import nekbench_amd as _nbx
    _nbx.lib()
    return _nbx


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    return torch.device("cuda:0")
