import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "gc-fg_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ora():
    import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def g():
    import gcfg

    gcfg.lib()
    return gcfg


def random_instance(rng, W, H, tmax=40, emax=20, edges=True):
    pos = rng.integers(0, tmax, (H, W)).astype(np.int32)
    neg = rng.integers(0, tmax, (H, W)).astype(np.int32)
    if not edges:
        return pos, neg, None
    E = tuple(rng.integers(0, emax, (H, W)).astype(np.int32) for _ in range(4))
    return pos, neg, E


def debug_kat():
    """Debug.cu:24-58: W=4, H=15, penalty 50."""
    W, H = 4, 15
    pos = np.zeros((H, W), np.int32)
    neg = np.zeros((H, W), np.int32)
    pos[1, 1], pos[2, 2], pos[13, 1] = 4, 2, 1
    neg[1, 2], neg[2, 1], neg[11, 2], neg[13, 2] = 1, 2, 1, 4
    return W, H, pos, neg


def glibc_rand_instance(seed, W=400, H=400):
    """Debug.cu:60-74: srand(seed); pos = rand()%300, neg = rand()%300 per pixel, row-major."""
    import ctypes

    libc = ctypes.CDLL("libc.so.6")
    libc.srand(seed)
    vals = np.array([libc.rand() % 300 for _ in range(2 * W * H)], dtype=np.int32)
    pos = vals[0::2].reshape(H, W).copy()
    neg = vals[1::2].reshape(H, W).copy()
    return pos, neg
