"""Initial centroids on the GPU (km_init_centroids_*): every method equals the
oracle's restatement bit for bit, from host and device memory, and the
preconditions follow kmeans_init (detail/compute_ops.hpp:40-56)."""

import numpy as np
import pytest

from conftest import gaussian_mixture

pytestmark = pytest.mark.gpu

CASES = [
    # n, d, k, seed
    (9, 2, 4, 0),
    (1000, 1, 7, 1),
    (5000, 3, 64, 2),
    (4097, 33, 130, 3),     # chunk boundary + 1
    (20000, 64, 200, 4),
    (12289, 128, 50, 5),
    (3000, 257, 16, 6),
]


@pytest.mark.parametrize("method", [0, 1, 2], ids=["dense", "random_dense", "plus_plus_dense"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("case", CASES, ids=["n%d_d%d_k%d" % c[:3] for c in CASES])
def test_init_matches_oracle(engine, oracle, method, dtype, case):
    n, d, k, seed = case
    x = gaussian_mixture(n, d, max(2, k // 3), seed=seed, dtype=dtype)
    want = oracle.init_centroids(x, k, method, seed * 7919 + 11)
    got = engine.init_centroids(x, k, method, seed * 7919 + 11)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_init_device_pointers_and_duplicates(engine, oracle):
    import torch

    rng = np.random.default_rng(3)
    x = rng.integers(0, 3, size=(6000, 5)).astype(np.float32)  # many duplicate rows
    for method in (0, 1, 2):
        want = oracle.init_centroids(x, 40, method, 99)
        got = engine.init_centroids(torch.from_numpy(x).cuda(), 40, method, 99)
        torch.cuda.synchronize()
        assert np.array_equal(got.cpu().numpy(), want), method
    same = np.full((300, 4), 2.5, np.float64)  # zero total D^2 after the first pick
    assert np.array_equal(engine.init_centroids(same, 5, 2, 1), oracle.init_centroids(same, 5, 2, 1))


def test_init_preconditions(engine):
    from onedal_amd.engine import DomainError, Unimplemented

    x = np.zeros((10, 3), np.float32)
    with pytest.raises(DomainError, match="Cluster count exceeds data row count"):
        engine.init_centroids(x, 10, 2)
    with pytest.raises(DomainError, match="lower than or equal to zero"):
        engine.init_centroids(x, 0, 0)
    with pytest.raises(Unimplemented, match="parallel dense"):
        engine.init_centroids(x, 3, 3)
    assert engine.init_centroids(x, 3, 0).shape == (3, 3)  # usable afterwards


def test_train_from_plus_plus_init(engine, oracle):
    """k-means++ centroids feed Lloyd training; the whole pipeline equals the
    oracle's (init restatement + canonical Lloyd)."""
    x = gaussian_mixture(30000, 16, 12, seed=8)
    c0 = engine.init_centroids(x, 12, 2, 777)
    assert np.array_equal(c0, oracle.init_centroids(x, 12, 2, 777))
    c, lab, obj, it = engine.train(x, c0, 10, 0.0)
    wc, wl, wo, wi, _ = oracle.canonical_lloyd(x, c0, 10, 0.0)
    assert np.array_equal(lab, wl) and np.array_equal(c, wc) and obj == wo
