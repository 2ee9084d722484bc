import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def ctx():
    import thea_amd
    return thea_amd.default_context(0)


def random_mesh(g, nt, spread=1.0, size=0.1):
    """nt independent random triangles (indexed, no shared vertices)."""
    c = g.uniform(-spread, spread, size=(nt, 1, 3))
    v = (c + g.normal(0, size, size=(nt, 3, 3))).reshape(-1, 3).astype(np.float32)
    t = np.arange(3 * nt, dtype=np.int32).reshape(-1, 3)
    return v, t
