import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity of the HIP path")


@pytest.fixture(scope="session")
def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(params=["0", "1"], ids=["plain-ranges", "xcd-banded"])
def xcd_layout(request, monkeypatch):
    """Both work-list layouts of a single case: plain [lo, hi) ranges (the
    default above 600k cells) and XCD-banded lists (the default below);
    RC_XCD is read when a case is created."""
    monkeypatch.setenv("RC_XCD", request.param)
    return request.param
