"""CPU tests: pin the oracle against the reference's own golden vectors and
against what the mathematics fixes.  No GPU needed."""

import json
import os

import numpy as np
import pytest

from conftest import gaussian_mixture

HERE = os.path.dirname(os.path.abspath(__file__))


def load_gold(dtype):
    with open(os.path.join(HERE, "golden", "gold_dataset.json")) as f:
        g = json.load(f)
    n, d, k = g["row_count"], g["column_count"], g["cluster_count"]
    x = np.array(g["data"], dtype).reshape(n, d)
    c0 = np.array(g["initial_centroids"], dtype).reshape(k, d)
    ec = np.array(g["expected_centroids"], dtype).reshape(k, d)
    er = np.array(g["expected_responses"], np.int32)
    return g, x, c0, ec, er


def match_map(expected, got):
    """test/fixture.hpp:677-706: nearest expected centroid for each computed one, must be a bijection."""
    k = expected.shape[0]
    mm = np.array([int(np.argmin(((expected - got[i]) ** 2).sum(1))) for i in range(k)])
    assert sorted(mm.tolist()) == list(range(k))
    return mm


def rel_close(a, b, tol):
    """test/fixture.hpp:538-558 style relative check: |l-r| / (|l|+|r|+tiny) <= tol."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.all(np.abs(a - b) / (np.abs(a) + np.abs(b) + 1e-300) <= tol)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("which", ["daal", "canonical"])
def test_gold_dataset(oracle, dtype, which):
    # test/fixture.hpp:156-203: centroids rel 1e-7 after matching, labels, objective rel 1e-7
    g, x, c0, ec, er = load_gold(dtype)
    if which == "daal":
        c, lab, obj, it = oracle.daal_lloyd(x, c0, g["max_iteration_count"], 0.0, stop_rule=0,
                                            block_size=64)
    else:
        c, lab, obj, it, _ = oracle.canonical_lloyd(x, c0, g["max_iteration_count"], 0.0)
    mm = match_map(ec, c)
    assert rel_close(ec[mm], c, 1e-7)
    assert np.array_equal(er, mm[lab])
    assert abs(obj - g["expected_objective"]) / g["expected_objective"] < 1e-7


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("which", ["daal", "canonical"])
def test_empty_clusters_fixture(oracle, dtype, which):
    # test/fixture.hpp:205-219: 1-D data, C0 = {-10,-10,-10}, one iteration
    x = np.array([-10, -9.5, -9, -8.5, -8, -1, 1, 9, 9.5, 10], dtype).reshape(10, 1)
    c0 = np.array([-10, -10, -10], dtype).reshape(3, 1)
    if which == "daal":
        c, lab, obj, it = oracle.daal_lloyd(x, c0, 1, 0.0, stop_rule=2, block_size=4)
    else:
        c, lab, obj, it, _ = oracle.canonical_lloyd(x, c0, 1, 0.0)
    assert rel_close(c.ravel(), np.array([-1.65, 10, 9.5], dtype), 5 * np.finfo(dtype).eps * 100)
    assert lab.tolist() == [0, 0, 0, 0, 0, 0, 0, 2, 2, 1]
    assert it == 1


def test_stress_identity(oracle):
    # test/fixture.hpp:289-308 (scaled down): C0 = X gives labels = iota, centroids = X
    rng = np.random.default_rng(7777)
    x = rng.uniform(-0.2, 0.5, size=(256, 64)).astype(np.float32)
    for fn in ("daal", "canonical"):
        if fn == "daal":
            c, lab, obj, it = oracle.daal_lloyd(x, x, 1, 0.0, stop_rule=2, block_size=32)
        else:
            c, lab, obj, it, _ = oracle.canonical_lloyd(x, x, 1, 0.0)
        assert np.array_equal(lab, np.arange(256))
        assert np.array_equal(c, x)
        assert obj == 0.0


def test_one_cluster_mean(oracle):
    # test/fixture.hpp:254-287 (scaled down): k = 1 -> centroid is the column mean
    rng = np.random.default_rng(7777)
    x = rng.uniform(-0.2, 0.5, size=(20000, 16)).astype(np.float32)
    c, lab, obj, it, _ = oracle.canonical_lloyd(x, x[:1], 1, 0.0)
    mean = x.astype(np.float64).mean(0)
    assert np.max(np.abs(c[0] - mean)) <= 1e-6 * np.max(np.abs(mean)) + 1e-7
    var_sum = ((x.astype(np.float64) - c[0].astype(np.float64)) ** 2).sum()
    assert abs(obj - var_sum) / var_sum < 1e-10


@pytest.mark.parametrize("n,k,d", [(10, 3, 2), (10001, 37, 33)])
def test_canonical_assign_is_direct_bruteforce(oracle, n, k, d):
    # backend/gpu/test/cluster_updater_dpc.cpp:288-333, 537-557: labels equal the
    # direct sum (x-c)^2 argmin with strict '<' (in F, mul then add)
    rng = np.random.default_rng(7777)
    c = rng.uniform(-1.0, 2.7, size=(k, d)).astype(np.float32)
    x = rng.uniform(-0.9, 1.7, size=(n, d)).astype(np.float32)
    lab, best = oracle.canonical_assign(x, c)
    # numpy restatement for the small case, pure loops
    if n <= 10:
        for i in range(n):
            bestv, bi = None, -1
            for j in range(k):
                acc = np.float32(0)
                for f in range(d):
                    diff = np.float32(x[i, f] - c[j, f])
                    acc = np.float32(acc + np.float32(diff * diff))
                if bestv is None or acc < bestv:
                    bestv, bi = acc, j
            assert lab[i] == bi and best[i] == bestv
    # the FP64 brute force agrees wherever the top-2 gap is resolvable in float
    l64, b1, b2 = oracle.fp64_assign(x, c)
    gap_ok = (b2 - b1) > 1e-5 * (b1 + 1.0)
    assert np.array_equal(lab[gap_ok], l64[gap_ok])
    assert gap_ok.mean() > 0.99


def test_canonical_sums_closed_form(oracle):
    # partial/merge reduction (cluster_updater_dpc.cpp:504-535): per-cluster sums
    rng = np.random.default_rng(7777)
    for d in (1, 3, 17, 33, 64, 100):
        x = rng.uniform(-0.9, 1.7, size=(10001, d)).astype(np.float32)
        lab = rng.integers(0, 37, size=10001).astype(np.int32)
        s1, counts = oracle.canonical_sums(x, lab, 37)
        ref = np.zeros((37, d))
        np.add.at(ref, lab, x.astype(np.float64))
        assert np.array_equal(counts, np.bincount(lab, minlength=37))
        assert np.allclose(s1, ref, rtol=1e-13, atol=1e-12)


def test_canonical_objective_matches_exact(oracle):
    rng = np.random.default_rng(7)
    x = rng.uniform(-5, 5, size=(30000, 7)).astype(np.float32)
    c = rng.uniform(-5, 5, size=(9, 7)).astype(np.float32)
    lab = rng.integers(0, 9, size=30000).astype(np.int32)
    obj = oracle.canonical_objective(x, c, lab)
    ref = ((x.astype(np.float64) - c[lab].astype(np.float64)) ** 2).sum()
    assert abs(obj - ref) / ref < 1e-12


def test_daal_vs_canonical_mixture(oracle):
    # the two restatements differ only by rounding: on a well-separated mixture
    # they must agree on labels and (to rounding) on centroids and objective
    x = gaussian_mixture(5000, 8, 6, seed=3, center_range=10.0)
    c0 = x[:6].copy()
    c1, l1, o1, i1 = oracle.daal_lloyd(x, c0, 15, 0.0, stop_rule=2, block_size=128)
    c2, l2, o2, i2, _ = oracle.canonical_lloyd(x, c0, 15, 0.0)
    assert i1 == i2 == 15
    assert (l1 == l2).mean() > 0.999
    assert np.allclose(c1, c2, rtol=1e-4, atol=1e-4)
    assert abs(o1 - o2) / o2 < 1e-5


def test_planted_mistake_is_caught(oracle):
    """The checks above are not tautologies: a perturbed oracle output fails them."""
    g, x, c0, ec, er = load_gold(np.float64)
    c, lab, obj, it, _ = oracle.canonical_lloyd(x, c0, 100, 0.0)
    bad = c.copy()
    bad[0, 0] *= 1.0 + 1e-6
    mm = match_map(ec, bad)
    assert not rel_close(ec[mm], bad, 1e-7)
    assert abs(obj * (1 + 1e-6) - g["expected_objective"]) / g["expected_objective"] > 1e-7
