import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def pytest_collection_modifyitems(config, items):
    # GPU tests print nothing for long stretches; make sure they are bounded.
    for item in items:
        if item.get_closest_marker("gpu") is not None and item.get_closest_marker("timeout") is None:
            item.add_marker(pytest.mark.timeout(300))


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.load_library()
    return oracle


def rel_disp_error(x_gpu, x_ref, x0):
    """max_i |x_gpu - x_ref| / max_i |x_ref - x0| (node-displacement error)."""
    den = np.linalg.norm(x_ref - x0, axis=1).max()
    num = np.linalg.norm(x_gpu - x_ref, axis=1).max()
    return num / den if den > 0 else num
