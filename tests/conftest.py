import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def engine():
    from onedal_amd.engine import Engine

    eng = Engine(0)
    yield eng
    eng.close()


def gaussian_mixture(n, d, k, seed, center_range=10.0, sigma=(1.0, 1.0), dtype=np.float32,
                     anisotropic=False):
    """Synthetic Gaussian-mixture rows (same recipe as the benchmark workloads)."""
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-center_range, center_range, size=(k, d))
    if anisotropic:
        scale = 10.0 ** rng.uniform(-1.0, 1.0, size=(k, d))
    else:
        scale = np.repeat(rng.uniform(sigma[0], sigma[1], size=(k, 1)), d, axis=1)
    lab = rng.integers(0, k, size=n)
    x = centers[lab] + rng.standard_normal((n, d)) * scale[lab]
    return x.astype(dtype)
