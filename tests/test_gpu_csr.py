"""GPU parity of the CSR Lloyd path (method lloyd_csr) with the oracle's CSR
restatement: labels, centroids, objective and iteration count bit for bit, on
the shapes of the reference's sparse tests (kmeans/test/batch.cpp:276-350:
5/16/32/128 clusters, 5% of entries stored), zero- and one-based, plus the edge
cases (empty rows, unsorted and duplicate columns, empty clusters, k = 1, no
iterations) and the structure checks."""
import numpy as np
import pytest

from test_oracle_csr import reorder, sparse_blobs

pytestmark = pytest.mark.gpu


def check_train(engine, oracle, x, c0, max_iter, thr):
    c, labels, obj, iters = engine.train_csr(x.values, x.cols, x.offs, x.d, c0, max_iter, thr,
                                             base=x.base)
    rc, rl, ro, ri = oracle.csr_lloyd(x, c0, max_iter, thr)
    assert iters == ri
    assert np.array_equal(labels, rl)
    assert np.array_equal(c.view(np.uint8), rc.view(np.uint8))
    assert obj == ro
    return c, labels, obj, iters


CASES = [
    # (n, d, k, density, iters, thr, dtype, base)  -- reference sparse test shapes
    (50, 20, 5, 0.05, 10, 0.01, np.float32, 0),
    (200, 100, 16, 0.05, 10, 0.01, np.float64, 1),
    (10000, 100, 32, 0.05, 30, 0.01, np.float32, 1),
    (100000, 200, 128, 0.05, 10, 0.01, np.float64, 0),
    # wider: several 64-entry groups per row, k above 128 (four centroid slots per lane)
    (20000, 1000, 300, 0.1, 5, 0.0, np.float32, 0),
    (5000, 33, 70, 0.5, 8, 0.0, np.float64, 1),
    (3000, 7, 3, 1.0, 6, 0.0, np.float32, 0),
    # four 16-byte vectors per lane (k above 512 floats / 256 doubles)
    (6000, 64, 600, 0.2, 4, 0.0, np.float32, 1),
    (4000, 50, 300, 0.2, 4, 0.0, np.float64, 0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "n%d_d%d_k%d_%s_b%d" % (
    c[0], c[1], c[2], np.dtype(c[6]).name, c[7]))
def test_train_matches_oracle(engine, oracle, case):
    n, d, k, dens, iters, thr, dtype, base = case
    print("csr train", case, flush=True)
    x, _ = sparse_blobs(n, d, k, dens, seed=n + d + k, dtype=dtype, base=base)
    rng = np.random.default_rng(k)
    c0 = x.to_dense()[rng.choice(n, k, replace=False)]
    check_train(engine, oracle, x, c0, iters, thr)


def test_infer_matches_oracle(engine, oracle):
    for dtype in (np.float32, np.float64):
        x, _ = sparse_blobs(30000, 300, 40, 0.05, seed=7, dtype=dtype)
        c = np.random.default_rng(1).normal(0.0, 3.0, size=(40, 300)).astype(dtype)
        labels, obj = engine.infer_csr(x.values, x.cols, x.offs, x.d, c, objective=True)
        rl, closest = oracle.csr_assign(x, c)
        assert np.array_equal(labels, rl)
        assert obj == oracle.csr_objective(closest)


def test_unsorted_duplicates_and_empty_rows(engine, oracle):
    x, _ = sparse_blobs(6000, 64, 10, 0.2, seed=31)
    x = reorder(x, seed=32, base=0)
    # empty rows and duplicated entries
    from oracle.oracle import Csr

    vals, cols, offs = [], [], [0]
    for i in range(x.n):
        a, b = x.offs[i], x.offs[i + 1]
        if i % 7 == 0:
            pass  # empty row
        else:
            v, cc = x.values[a:b], x.cols[a:b]
            if i % 5 == 0 and b > a:
                v = np.concatenate([v, v[:1] * 0.5])
                cc = np.concatenate([cc, cc[:1]])
            vals.append(v)
            cols.append(cc)
        offs.append(offs[-1] + (sum(len(q) for q in vals) - offs[-1]))
    y = Csr(np.concatenate(vals).astype(np.float32), np.concatenate(cols), np.array(offs), x.d, 0)
    c0 = y.to_dense()[1:11].copy()
    check_train(engine, oracle, y, c0, 6, 0.0)


def test_empty_clusters(engine, oracle):
    x, _ = sparse_blobs(4000, 50, 6, 0.2, seed=41, dtype=np.float64)
    c0 = np.vstack([x.to_dense()[:6], np.full((4, 50), 1e3)])  # four far centroids: empty
    c, labels, obj, iters = check_train(engine, oracle, x, c0, 5, 0.0)
    assert engine.stats()["empty_cluster_events"] >= 1


def test_small_edges(engine, oracle):
    x, _ = sparse_blobs(500, 20, 4, 0.2, seed=51)
    c0 = x.to_dense()[:1].copy()
    check_train(engine, oracle, x, c0, 3, 0.0)                 # k = 1
    c0 = x.to_dense()[:4].copy()
    c, labels, obj, iters = check_train(engine, oracle, x, c0, 0, 0.0)  # no iterations
    assert iters == 0 and np.array_equal(c, c0)
    c0 = x.to_dense()[:500].copy()
    check_train(engine, oracle, x, c0, 2, 0.0)                 # k = n


def test_repeatable_and_device_pointers(engine, oracle):
    import torch

    x, _ = sparse_blobs(20000, 120, 24, 0.05, seed=61)
    c0 = x.to_dense()[:24].copy()
    a = engine.train_csr(x.values, x.cols, x.offs, x.d, c0, 8)
    b = engine.train_csr(x.values, x.cols, x.offs, x.d, c0, 8)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
    dv = torch.from_numpy(x.values).cuda()
    dc = torch.from_numpy(x.cols).cuda()
    do = torch.from_numpy(x.offs).cuda()
    dc0 = torch.from_numpy(c0).cuda()
    c, labels, obj, iters = engine.train_csr(dv, dc, do, x.d, dc0, 8)
    assert np.array_equal(c.cpu().numpy(), a[0])
    assert np.array_equal(labels.cpu().numpy(), a[1])
    assert obj == a[2] and iters == a[3]
    lab, o = engine.infer_csr(dv, dc, do, x.d, c, objective=True)
    ref_l, ref_o = engine.infer_csr(x.values, x.cols, x.offs, x.d, a[0], objective=True)
    assert np.array_equal(lab.cpu().numpy(), ref_l) and o == ref_o


def test_structure_checks(engine):
    from onedal_amd.engine import DomainError, InvalidArgument

    vals = np.ones(4, np.float32)
    c0 = np.zeros((1, 3), np.float32)
    with pytest.raises(InvalidArgument, match="Column indices"):
        engine.train_csr(vals, np.array([0, 1, 3, 2]), np.array([0, 2, 4]), 3, c0, 1)
    with pytest.raises(InvalidArgument, match="Row offsets"):
        engine.train_csr(vals, np.array([0, 1, 2, 2]), np.array([0, 3, 2]), 3, c0, 1)
    with pytest.raises(InvalidArgument, match="Row offsets"):
        engine.train_csr(vals, np.array([0, 1, 2, 2]), np.array([0, 2, 3]), 3, c0, 1)
    with pytest.raises(InvalidArgument, match="Column indices"):
        # one-based storage with a zero index
        engine.train_csr(vals, np.array([0, 1, 2, 2]), np.array([1, 3, 5]), 3, c0, 1, base=1)
    with pytest.raises(InvalidArgument, match="indexing"):
        engine.train_csr(vals, np.array([0, 1, 2, 2]), np.array([0, 2, 4]), 3, c0, 1, base=2)
    with pytest.raises(InvalidArgument, match="exceeds"):
        engine.train_csr(vals, np.array([0, 1, 2, 2]), np.array([0, 2, 4]), 3,
                         np.zeros((3, 3), np.float32), 1)
    with pytest.raises(DomainError, match="Cluster count"):
        engine.infer_csr(vals, np.array([0, 1, 2, 2]), np.array([0, 2, 4]), 3,
                         np.zeros((0, 3), np.float32))
    # the context stays usable
    labels, _ = engine.infer_csr(vals, np.array([0, 1, 2, 2]), np.array([0, 2, 4]), 3, c0)
    assert labels.tolist() == [0, 0]


def test_nan_semantics(engine, oracle):
    """NaN distances follow the sequential scan "best = s_0; if s_j < best": a NaN
    s_0 keeps label 0, a later NaN never wins, a NaN row value gives label 0."""
    from oracle.oracle import Csr

    x, _ = sparse_blobs(800, 24, 4, 0.3, seed=71, dtype=np.float64)
    used = int(x.cols[0])  # a column that row 0 (and its cluster) stores
    for j_nan in (0, 2):
        c = x.to_dense()[:4].copy()
        c[j_nan, used] = np.nan
        labels, _ = engine.infer_csr(x.values, x.cols, x.offs, x.d, c)
        rl, _ = oracle.csr_assign(x, c)
        assert np.array_equal(labels, rl)
    vals = x.values.copy()
    vals[x.offs[3]] = np.nan  # row 3's first entry
    y = Csr(vals, x.cols, x.offs, x.d, 0)
    c = x.to_dense()[:4].copy()
    labels, _ = engine.infer_csr(y.values, y.cols, y.offs, y.d, c)
    rl, _ = oracle.csr_assign(y, c)
    assert np.array_equal(labels, rl) and labels[3] == 0


def test_mixed_row_lengths(engine, oracle):
    """Rows of 0 to 100 stored entries in the same clusters: member groups of the
    chunk sums switch between the one-lane-per-entry path (rows of at most 64
    entries) and the general path (longer rows) inside one chunk."""
    from oracle.oracle import Csr

    x, _ = sparse_blobs(12000, 100, 6, 1.0, seed=81)
    rng = np.random.default_rng(82)
    keep = rng.integers(0, 101, size=x.n)
    vals, cols, offs = [], [], [0]
    for i in range(x.n):
        a = x.offs[i]
        b = a + min(int(keep[i]), int(x.offs[i + 1] - a))
        vals.append(x.values[a:b])
        cols.append(x.cols[a:b])
        offs.append(offs[-1] + (b - a))
    y = Csr(np.concatenate(vals), np.concatenate(cols), np.array(offs), x.d, 0)
    y = reorder(y, seed=83, base=1)
    c0 = y.to_dense()[rng.choice(x.n, 6, replace=False)].copy()
    check_train(engine, oracle, y, c0, 5, 0.0)
