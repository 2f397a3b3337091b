import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through librb.so's HIP kernels)")


def random_triangles(n, seed=0, spread=5.0, size=0.4):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n, 1, 3))
    v = (c + rng.normal(0, size, (n, 3, 3))).astype(np.float32).reshape(-1, 3)
    idx = np.arange(n * 3, dtype=np.uint32).reshape(-1, 3)
    return v, idx


def random_rays(n, seed=1, box=6.0, tmin=0.0, tmax=3.0e38):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-box, box, (n, 3))
    d = rng.normal(0, 1, (n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def rbm():
    import renderbox2_amd
    renderbox2_amd.lib()
    return renderbox2_amd


@pytest.fixture(scope="session")
def gpu(rbm):
    """Device handle for @pytest.mark.gpu tests: fails loudly without a GPU."""
    rbm.device_init(0)
    return rbm
