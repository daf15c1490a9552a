import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.lib()
    yield pyoracle
    pyoracle.pool_shutdown()


def announce(msg):
    """name the case on stdout before a GPU step (a fault ends the process)"""
    print("\n[gpu-case] " + msg, flush=True)
