import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libfvgpu.so")


def rel_l2(a, b):
    """Relative L2 difference ||a-b|| / max(||b||, tiny)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


@pytest.fixture(scope="session")
def built():
    """Make sure the host library and the oracle are built (CPU only)."""
    from myfvcomcode_amd import build
    import oracle
    if not os.path.exists(os.path.join(ROOT, "myfvcomcode_amd", "lib", "libfvgrid.so")):
        build.build_grid()
    oracle.build()
    return True
