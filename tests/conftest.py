"""pytest configuration: repository root on sys.path and the `gpu` marker.

Tests marked gpu need a HIP device and the built libkgmt.so; everything else
runs on the CPU (oracle checks, host logic, library loading)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU and the built libkgmt.so")


def pytest_runtest_setup(item):
    """PyTorch-ROCm brings its own HIP runtime; it must initialise the GPU
    before libkgmt's runtime does, or it reports no GPU afterwards (the tests
    use torch for device buffers), so GPU tests start torch's context first."""
    if item.get_closest_marker("gpu") is not None:
        try:
            import torch
        except ImportError:
            return
        torch.cuda.is_available()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    O.lib()
    return O
