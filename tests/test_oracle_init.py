"""CPU checks that pin the oracle's initial-centroid methods (kmeans_init):
the reference's own property test, a published known answer of the generator,
and the sampling laws each method must follow.  The reference's draws come from
DAAL's mt19937 and cannot be reproduced here, so the random methods are pinned
by their distributions."""

import math

import numpy as np
import pytest


def test_splitmix64_known_answer(oracle):
    """First output of SplitMix64 seeded with 0 (Steele, Lea, Flood 2014;
    the reference implementation's published value)."""
    L = oracle.lib()
    assert L.oracle_init_mix(0) == 0xE220A8397B1DCDAF
    assert L.oracle_init_mix(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


@pytest.mark.parametrize("method", [0, 1, 2])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_reference_fixture_distinct_rows(oracle, method, dtype):
    """kmeans_init/test/batch.cpp:40-56 with the check of fixture.hpp:61-90: the
    k centroids are k distinct rows of the data (every seed)."""
    data = np.array([1.0, 1.0, 2.0, 2.0, 1.0, 2.0, 2.0, 1.0, -1.0, -1.0, -1.0, -2.0, -2.0, -1.0, -2.0,
                     -2.0], dtype=dtype).reshape(8, 2)
    for seed in range(64):
        c = oracle.init_centroids(data, 4, method, seed)
        used = set()
        for row in c:
            match = [j for j in range(8) if np.array_equal(data[j], row) and j not in used]
            assert match, (method, seed, c)
            used.add(match[0])
        assert len(used) == 4


def test_dense_is_first_rows(oracle):
    x = np.random.default_rng(0).standard_normal((50, 7)).astype(np.float32)
    assert np.array_equal(oracle.init_centroids(x, 9, 0, 123), x[:9])


def test_random_dense_is_uniform_without_replacement(oracle):
    """Each of n rows is picked with probability k/n; picks are distinct."""
    n, k, trials = 10, 3, 3000
    x = np.arange(n, dtype=np.float64).reshape(n, 1)
    hits = np.zeros(n)
    for seed in range(trials):
        c = oracle.init_centroids(x, k, 1, seed).ravel().astype(int)
        assert len(set(c)) == k
        hits[c] += 1
    p = k / n
    sigma = math.sqrt(trials * p * (1 - p))
    assert np.all(np.abs(hits - trials * p) < 5 * sigma), hits


def test_plus_plus_follows_the_d2_law(oracle):
    """Rows at 0, 1 and 10 on a line, k = 2: the first pick is uniform and the
    second follows P(j | i) = D(i, j)^2 / sum_j D(i, j)^2.  Sampling by D instead
    of D^2, or uniformly, fails this."""
    x = np.array([[0.0], [1.0], [10.0]])
    trials = 6000
    first = np.zeros(3)
    pair = np.zeros((3, 3))
    for seed in range(trials):
        c = oracle.init_centroids(x, 2, 2, seed).ravel()
        i, j = int(np.flatnonzero(x.ravel() == c[0])[0]), int(np.flatnonzero(x.ravel() == c[1])[0])
        assert i != j
        first[i] += 1
        pair[i, j] += 1
    sig = math.sqrt(trials * (1 / 3) * (2 / 3))
    assert np.all(np.abs(first - trials / 3) < 5 * sig), first
    pts = x.ravel()
    for i in range(3):
        w = (pts - pts[i]) ** 2
        p = w / w.sum()
        m = first[i]
        for j in range(3):
            s = math.sqrt(max(m * p[j] * (1 - p[j]), 1.0))
            assert abs(pair[i, j] - m * p[j]) < 5 * s + 1, (i, j, pair[i], m * p)


def test_plus_plus_skips_duplicates_and_handles_zero_total(oracle):
    """Rows equal to a chosen centroid have D^2 = 0 and are never drawn; when every
    row equals a chosen centroid the result is still rows of the data."""
    x = np.zeros((100, 3), np.float32)
    x[57] = [1.0, 2.0, 3.0]
    for seed in range(20):
        c = oracle.init_centroids(x, 2, 2, seed)
        assert {tuple(c[0]), tuple(c[1])} == {(0.0, 0.0, 0.0), (1.0, 2.0, 3.0)}
    same = np.ones((20, 4))
    c = oracle.init_centroids(same, 3, 2, 5)
    assert np.array_equal(c, np.ones((3, 4)))
