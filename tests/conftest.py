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


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    O.lib()
    return O
