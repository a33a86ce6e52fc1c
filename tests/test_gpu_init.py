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


# -- CSR rows (km_init_centroids_csr_*) ----------------------------------------------


def _sparse_rows(n, d, density, seed, dtype, shuffle=False, dup=False):
    """CSR arrays (zero-based) of n rows: some empty, some with an explicit 0 or
    -0 stored; shuffle permutes each row's entries, dup stores some columns twice."""
    rng = np.random.default_rng(seed)
    vals, cols, offs = [], [], [0]
    for i in range(n):
        m = 0 if i % 17 == 3 else int(rng.binomial(d, density))
        c = np.sort(rng.choice(d, size=m, replace=False)).astype(np.int64)
        v = (rng.normal(size=m) * 3.0 + (i % 5) * 4.0).astype(dtype)
        if m and i % 11 == 0:
            v[0] = dtype(0.0)
        if m and i % 13 == 0:
            v[-1] = dtype(-0.0)
        if dup and m:
            extra = rng.choice(c, size=max(1, m // 3))
            c = np.concatenate([c, extra])
            v = np.concatenate([v, rng.normal(size=extra.size).astype(dtype)])
        if shuffle and c.size > 1:
            p = rng.permutation(c.size)
            c, v = c[p], v[p]
        vals.append(v)
        cols.append(c)
        offs.append(offs[-1] + c.size)
    return (np.concatenate(vals).astype(dtype), np.concatenate(cols).astype(np.int64),
            np.asarray(offs, np.int64))


def _expand(vals, cols, offs, d):
    """The dense rows: each stored value added onto a zero, in stored order."""
    n = offs.size - 1
    x = np.zeros((n, d), vals.dtype)
    for i in range(n):
        for p in range(offs[i], offs[i + 1]):
            x[i, cols[p]] = x[i, cols[p]] + vals[p]
    return x


CSR_CASES = [
    # n, d, k, density, seed
    (300, 3, 5, 0.5, 1),
    (5000, 40, 16, 0.2, 2),
    (4097, 1100, 8, 0.01, 3),   # features beyond one staged pass
    (2000, 64, 130, 0.05, 4),
]


@pytest.mark.parametrize("method", [0, 1, 2], ids=["dense", "random_dense", "plus_plus_dense"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("case", CSR_CASES, ids=["n%d_d%d_k%d" % c[:3] for c in CSR_CASES])
def test_init_csr_equals_dense_expansion(engine, oracle, method, dtype, case):
    n, d, k, density, seed = case
    vals, cols, offs = _sparse_rows(n, d, density, seed, dtype)
    x = _expand(vals, cols, offs, d)
    want = oracle.init_centroids(x, k, method, seed * 31 + 5)
    got = engine.init_centroids_csr(vals, cols, offs, d, k, method, seed * 31 + 5)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    # one-based indexing
    got1 = engine.init_centroids_csr(vals, cols + 1, offs + 1, d, k, method, seed * 31 + 5, base=1)
    assert np.array_equal(got1.view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_init_csr_unsorted_repeated_columns_device(engine, oracle, dtype):
    import torch

    d, k = 70, 12
    vals, cols, offs = _sparse_rows(3000, d, 0.1, 9, dtype, shuffle=True, dup=True)
    x = _expand(vals, cols, offs, d)
    for method in (0, 1, 2):
        want = oracle.init_centroids(x, k, method, 123)
        got = engine.init_centroids_csr(vals, cols, offs, d, k, method, 123)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), method
        dv, dc, do = (torch.from_numpy(a).cuda() for a in (vals, cols, offs))
        gd = engine.init_centroids_csr(dv, dc, do, d, k, method, 123)
        torch.cuda.synchronize()
        assert np.array_equal(gd.cpu().numpy().view(np.uint8), want.view(np.uint8)), method
    # sorted rows from device memory take the in-place walk
    vals, cols, offs = _sparse_rows(3000, d, 0.1, 10, dtype)
    want = oracle.init_centroids(_expand(vals, cols, offs, d), k, 2, 7)
    gd = engine.init_centroids_csr(*(torch.from_numpy(a).cuda() for a in (vals, cols, offs)), d, k, 2, 7)
    torch.cuda.synchronize()
    assert np.array_equal(gd.cpu().numpy().view(np.uint8), want.view(np.uint8))


def test_init_csr_preconditions(engine):
    from onedal_amd.engine import DomainError, InvalidArgument, Unimplemented

    vals, cols, offs = _sparse_rows(10, 4, 0.5, 0, np.float32)
    with pytest.raises(DomainError, match="Cluster count exceeds data row count"):
        engine.init_centroids_csr(vals, cols, offs, 4, 10, 2)
    with pytest.raises(Unimplemented, match="parallel dense"):
        engine.init_centroids_csr(vals, cols, offs, 4, 3, 3)
    bad = cols.copy()
    bad[0] = 4
    with pytest.raises(InvalidArgument, match="Column indices"):
        engine.init_centroids_csr(vals, bad, offs, 4, 3, 2)
    assert engine.init_centroids_csr(vals, cols, offs, 4, 3, 2).shape == (3, 4)


def test_init_csr_edge_shapes(engine, oracle):
    """k = 1, d = 1, every row empty (all D^2 zero after the first pick), a
    single stored entry, and nnz = 0 with device arrays."""
    import torch

    cases = []
    # every row empty: the expansion is all zeros
    cases.append((np.zeros(0, np.float32), np.zeros(0, np.int64), np.zeros(51, np.int64), 7, 5))
    # one stored entry in the whole matrix
    cases.append((np.array([3.5], np.float64), np.array([2], np.int64),
                  np.array([0] * 20 + [1] * 31, np.int64), 4, 3))
    # d = 1, k = 1
    v = np.arange(1, 41, dtype=np.float32)
    cases.append((v, np.zeros(40, np.int64), np.arange(41, dtype=np.int64), 1, 1))
    for vals, cols, offs, d, k in cases:
        x = _expand(vals, cols, offs, d)
        for method in (0, 1, 2):
            want = oracle.init_centroids(x, k, method, 5)
            got = engine.init_centroids_csr(vals, cols, offs, d, k, method, 5)
            assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (d, k, method)
    vals, cols, offs, d, k = cases[0]
    gd = engine.init_centroids_csr(torch.from_numpy(vals).cuda(), torch.from_numpy(cols).cuda(),
                                   torch.from_numpy(offs).cuda(), d, k, 2, 5)
    torch.cuda.synchronize()
    assert np.array_equal(gd.cpu().numpy(), oracle.init_centroids(_expand(vals, cols, offs, d), k, 2, 5))
