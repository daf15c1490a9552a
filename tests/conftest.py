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


def oracle_view(oracle, sc, clear_bgra):
    """oracle render of one view into texture-layout targets (rows bottom-up,
    tight scanlines, every row cleared), as psgpu_clear_target leaves them"""
    import numpy as np
    f = oracle.Frame(sc.width, sc.height, depth_scanline=sc.width, top_down=False)
    f.colour.view(np.uint32)[:] = clear_bgra
    f.clear_depth(1.0)
    return oracle.render_scene(sc, frame=f, clear=False)
