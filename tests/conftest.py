"""Shared fixtures.  GPU tests carry ``@pytest.mark.gpu`` and run on a ROCm device."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X / gfx950)")


@pytest.fixture(scope="session")
def device():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("GPU test selected but no ROCm device is visible")
    torch.cuda.set_device(0)
    return torch.device("cuda:0")


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case on stdout before any GPU work so a crash is attributable
    if request.node.get_closest_marker("gpu") is not None:
        print(f"\n[gpu-case] {request.node.nodeid}", flush=True)
    yield
