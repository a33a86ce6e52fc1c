"""CPU checks of the oracle's CSR restatement (method lloyd_csr).

What pins it:
  * the reference's own known answer for CSR k-means
    (cpp/oneapi/dal/algo/kmeans/test/batch.cpp:404-478: three rows, two
    clusters, objective 100, centroids equal to the dense run's), zero- and
    one-based;
  * the dense oracle (pinned to the reference's golden vectors elsewhere): on the
    same rows and the same labels the CSR centroid sums are the dense ones bit for
    bit, so a CSR run and a dense run from the same start agree in labels and
    centroids exactly on the reference's k-means example data
    (data/kmeans_csr_train_data.csv is the dense example data stored as CSR);
  * mathematics: labels are the exact argmin of |x - c|^2 on separated data, the
    closest distance equals |x - c_label|^2 up to rounding, stored column order
    and zero-based/one-based storage do not change the result, and an empty
    cluster takes the farthest row.
"""
import os

import numpy as np


def sparse_blobs(n, d, k, density, seed, dtype=np.float32, base=0):
    """Rows of k clusters; cluster j stores entries only on its own column
    support (so clusters stay apart), values = center + noise."""
    from oracle.oracle import Csr

    rng = np.random.default_rng(seed)
    m = max(1, int(round(density * d)))
    supports = [rng.choice(d, size=m, replace=False) for _ in range(k)]
    centers = rng.uniform(2.0, 10.0, size=(k, d))
    truth = rng.integers(0, k, size=n)
    vals, cols, offs = [], [], [0]
    for i in range(n):
        j = truth[i]
        cs = np.sort(supports[j])
        v = centers[j, cs] + rng.normal(0.0, 0.3, size=cs.shape[0])
        vals.append(v)
        cols.append(cs)
        offs.append(offs[-1] + cs.shape[0])
    x = Csr(np.concatenate(vals).astype(dtype), np.concatenate(cols).astype(np.int64) + base,
            np.array(offs, np.int64) + base, d, base)
    return x, truth


def reorder(x, seed, base):
    """Same rows with each row's entries in a random stored order, stored with base."""
    from oracle.oracle import Csr

    rng = np.random.default_rng(seed)
    vals, cols = x.values.copy(), x.cols.copy() - x.base
    offs = x.offs - x.base
    for i in range(x.n):
        a, b = offs[i], offs[i + 1]
        p = rng.permutation(b - a) + a
        vals[a:b], cols[a:b] = vals[p], cols[p]
    return Csr(vals, cols + base, offs + base, x.d, base)


def test_reference_known_answer(oracle):
    """batch.cpp:404-478: x = diag(10, 10, 20), C0 = [(5,5,0), (0,0,20)], one
    iteration: responses [0, 0, 1], objective 100 (the dense value), centroids
    equal to the dense run's."""
    for dtype in (np.float32, np.float64):
        for base in (0, 1):
            x = oracle.Csr(np.array([10, 10, 20], dtype), np.array([0, 1, 2]) + base,
                           np.array([0, 1, 2, 3]) + base, 3, base)
            c0 = np.array([[5, 5, 0], [0, 0, 20]], dtype)
            c, labels, obj, iters = oracle.csr_lloyd(x, c0, 1, 0.0)
            assert labels.tolist() == [0, 0, 1]
            assert obj == 100.0
            assert iters == 1
            np.testing.assert_array_equal(c, c0)
            dense = np.diag(np.array([10, 10, 20], dtype))
            cd, ld, objd, _, _ = oracle.canonical_lloyd(dense, c0, 1, 0.0)
            np.testing.assert_array_equal(c, cd)
            assert labels.tolist() == ld.tolist()
            assert objd == 100.0


def test_reference_example_data_matches_dense(oracle):
    """The reference's CSR example data (the dense example rows stored as CSR,
    one-based): CSR and dense runs from the same start give the same labels and
    bit-identical centroids (same members, same chunk sums)."""
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "kmeans_dense_example.npz"))
    x, c0 = z["data"], z["initial_centroids"]
    for dtype in (np.float32, np.float64):
        xd = x.astype(dtype)
        xc = oracle.Csr.from_dense(xd, base=1)
        assert np.array_equal(xc.to_dense(), xd)
        c, labels, obj, iters = oracle.csr_lloyd(xc, c0.astype(dtype), 5, 0.0)
        cd, ld, objd, itd, _ = oracle.canonical_lloyd(xd, c0.astype(dtype), 5, 0.0)
        assert np.array_equal(labels, ld)
        assert np.array_equal(c, cd)
        assert iters == itd == 5
        assert abs(obj - objd) <= 1e-4 * abs(objd)


def test_labels_are_exact_argmin(oracle):
    x, _ = sparse_blobs(3000, 40, 12, 0.2, seed=3)
    xd = x.to_dense()
    rng = np.random.default_rng(5)
    c = xd[rng.choice(x.n, 12, replace=False)]
    labels, closest = oracle.csr_assign(x, c)
    d2 = ((xd.astype(np.float64)[:, None, :] - c.astype(np.float64)[None]) ** 2).sum(-1)
    exact = d2.argmin(1)
    # rows whose best two exact distances are apart beyond rounding agree
    srt = np.sort(d2, axis=1)
    clear = (srt[:, 1] - srt[:, 0]) > 1e-3 * (1.0 + srt[:, 0])
    assert clear.mean() > 0.99
    assert np.array_equal(labels[clear], exact[clear])
    best = d2[np.arange(x.n), labels]
    np.testing.assert_allclose(closest.astype(np.float64), best, rtol=1e-4, atol=1e-3)


def test_column_order_and_indexing_do_not_matter(oracle):
    xa, _ = sparse_blobs(2000, 30, 8, 0.3, seed=11)
    xb = reorder(xa, seed=12, base=1)
    assert np.array_equal(xa.to_dense(), xb.to_dense())
    c0 = xa.to_dense()[:8].copy()
    ca, la, oa, ia = oracle.csr_lloyd(xa, c0, 10, 0.0)
    cb, lb, ob, ib = oracle.csr_lloyd(xb, c0, 10, 0.0)
    assert np.array_equal(la, lb)
    # centroid sums add stored entries per column in member order: the column
    # order inside a row does not change them
    assert np.array_equal(ca, cb)
    assert ia == ib
    assert abs(oa - ob) <= 1e-5 * abs(oa)


def test_duplicates_add_in_sums(oracle):
    """Duplicate entries of a row are summed (to_dense semantics) in the centroid."""
    x = oracle.Csr(np.array([1.0, 2.0, 4.0, 8.0], np.float64), np.array([0, 0, 1, 1]),
                   np.array([0, 2, 4]), 2)
    c0 = np.array([[0.0, 0.0]])
    c, labels, obj, iters = oracle.csr_lloyd(x, c0, 1, 0.0)
    np.testing.assert_array_equal(c, [[1.5, 6.0]])


def test_empty_cluster_takes_farthest_row(oracle):
    """Cluster 1 starts far away and gets no rows; it takes the row with the
    largest closest distance, densified, and the iteration objective drops by it."""
    xd = np.array([[1.0, 0.0], [0.0, 1.0], [0.0, 5.0]], np.float64)
    x = oracle.Csr.from_dense(xd)
    c0 = np.array([[0.0, 0.0], [100.0, 100.0]])
    c, labels, obj, iters = oracle.csr_lloyd(x, c0, 1, 0.0)
    np.testing.assert_array_equal(c[1], [0.0, 5.0])
    np.testing.assert_array_equal(c[0], [1.0 / 3.0, 2.0])
    # final assignment with the new centroids
    assert labels.tolist() == [0, 0, 1]


def test_stop_rule(oracle):
    x, _ = sparse_blobs(4000, 50, 10, 0.1, seed=21)
    c0 = x.to_dense()[:10].copy()
    _, _, _, it_all = oracle.csr_lloyd(x, c0, 30, 0.0)
    _, _, _, it_thr = oracle.csr_lloyd(x, c0, 30, 1e30)
    assert it_all == 30
    assert it_thr == 2  # obj_1 + huge > obj_0 stops after the second iteration
