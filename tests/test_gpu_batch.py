"""Batched training (km_batch_train_f32) on the GPU: every instance of a batch must
equal the oracle's canonical Lloyd on that instance alone, bit for bit (labels,
centroid bytes, objective), in eager and graph-replay mode."""

import numpy as np
import pytest

from conftest import gaussian_mixture

pytestmark = pytest.mark.gpu


def make_batch(seed, count):
    rng = np.random.default_rng(seed)
    dims = [1, 2, 3, 5, 8, 13, 16, 17, 32, 33, 48, 64, 100, 128]
    inst = []
    for i in range(count):
        d = dims[i % len(dims)]
        n = int(rng.integers(300, 9000))
        k = int(rng.integers(1, 131))
        k = min(k, n)
        x = gaussian_mixture(n, d, max(1, k // 2 + 1), seed=seed * 1000 + i, center_range=10.0,
                             sigma=(0.5, 2.0))
        if i % 5 == 0:
            # initial centroids with repeats: clusters that start empty
            idx = rng.choice(n, size=max(1, k // 2), replace=False)
            c0 = x[np.resize(idx, k)].copy()
        else:
            c0 = x[rng.choice(n, size=k, replace=False)].copy()
        inst.append((x, c0))
    # edge shapes: one cluster, n == k, a single column
    x = gaussian_mixture(500, 7, 3, seed=seed + 1)
    inst.append((x, x[:1].copy()))
    x = gaussian_mixture(64, 5, 4, seed=seed + 2)
    inst.append((x, x.copy()))
    x = gaussian_mixture(2000, 1, 4, seed=seed + 3)
    inst.append((x, x[:9].copy()))
    return inst


def run_batch(engine, inst, iters, graph):
    import torch

    xs = [torch.from_numpy(x).cuda() for x, _ in inst]
    cs = [torch.from_numpy(c).cuda() for _, c in inst]
    c_out, l_out, obj, its = engine.train_batch_device(xs, cs, iters, graph=graph)
    torch.cuda.synchronize()
    return ([c.cpu().numpy() for c in c_out], [l.cpu().numpy() for l in l_out], obj, its)


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_batch_bit_exact_vs_canonical(engine, oracle, graph):
    iters = 6
    inst = make_batch(11, 40)
    c_out, l_out, obj, its = run_batch(engine, inst, iters, graph)
    for i, (x, c0) in enumerate(inst):
        wc, wl, wo, wi = oracle.canonical_lloyd(x, c0, iters, 0.0)[:4]
        assert its[i] == wi == iters
        bad = np.flatnonzero(l_out[i] != wl)
        assert bad.size == 0, "instance %d (n=%d d=%d k=%d): labels differ at %s" % (
            i, x.shape[0], x.shape[1], c0.shape[0], bad[:10])
        assert np.array_equal(c_out[i].view(np.uint8), wc.view(np.uint8)), "instance %d centroids" % i
        assert obj[i] == wo, (i, obj[i], wo)


def test_batch_graph_replays_match_eager(engine):
    """The captured iteration replayed max_iter times gives the eager result, and a
    second call (fresh capture) gives it again."""
    inst = make_batch(23, 24)
    a = run_batch(engine, inst, 5, graph=False)
    b = run_batch(engine, inst, 5, graph=True)
    c = run_batch(engine, inst, 5, graph=True)
    for i in range(len(inst)):
        assert np.array_equal(a[1][i], b[1][i]) and np.array_equal(a[1][i], c[1][i])
        assert np.array_equal(a[0][i].view(np.uint8), b[0][i].view(np.uint8))
        assert np.array_equal(a[0][i].view(np.uint8), c[0][i].view(np.uint8))
        assert a[2][i] == b[2][i] == c[2][i]


def test_batch_matches_single_instance_engine(engine):
    """Batched and one-instance-per-call training agree (zero iterations included)."""
    import torch

    inst = make_batch(5, 12)
    for iters in (0, 3):
        c_out, l_out, obj, its = run_batch(engine, inst, iters, graph=False)
        for i, (x, c0) in enumerate(inst):
            c, lab, o, it = engine.train(x, c0, iters, 0.0)
            assert np.array_equal(lab, l_out[i]), i
            assert np.array_equal(c.view(np.uint8), c_out[i].view(np.uint8)), i
            assert o == obj[i] and it == its[i]
    torch.cuda.synchronize()


def test_batch_badargs(engine):
    import torch

    from onedal_amd.engine import DomainError, InvalidArgument

    x = torch.zeros((10, 3), dtype=torch.float32, device="cuda")
    c0 = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(InvalidArgument):
        engine.train_batch_device([x[:3]], [c0], 2)  # k > n
    with pytest.raises(DomainError):
        engine.train_batch_device([x], [c0], -1)
    # usable afterwards
    c, l, o, it = engine.train_batch_device([x], [c0], 1)
    assert it == [1]


def test_batch_every_tile_variant_identical(engine, oracle):
    """Every tile shape of the batched assignment gives the canonical result."""
    from onedal_amd.engine import KM_BATCH_TILE_VARIANTS, KM_OPT_BATCH_TILE

    inst = make_batch(31, 20)
    want = [oracle.canonical_lloyd(x, c0, 2, 0.0)[:3] for x, c0 in inst]
    try:
        for v in range(KM_BATCH_TILE_VARIANTS):
            engine.set_option(KM_OPT_BATCH_TILE, v + 1)
            c_out, l_out, obj, its = run_batch(engine, inst, 2, graph=(v % 2 == 1))
            for i, (wc, wl, wo) in enumerate(want):
                assert np.array_equal(l_out[i], wl), (v, i)
                assert np.array_equal(c_out[i].view(np.uint8), wc.view(np.uint8)), (v, i)
                assert obj[i] == wo, (v, i)
    finally:
        engine.set_option(KM_OPT_BATCH_TILE, 0)


def test_batch_autotune_table_and_results(oracle):
    """The autotuner fills every bucket with a valid variant, and batched training
    on the tuned context still gives the canonical result."""
    from onedal_amd.engine import KM_BATCH_TILE_BUCKETS, KM_BATCH_TILE_VARIANTS, Engine

    with Engine(0) as eng:
        table = eng.batch_autotune()
        assert len(table) == KM_BATCH_TILE_BUCKETS
        assert all(0 <= v < KM_BATCH_TILE_VARIANTS for v in table), table
        inst = make_batch(37, 16)
        c_out, l_out, obj, its = run_batch(eng, inst, 3, graph=True)
        for i, (x, c0) in enumerate(inst):
            wc, wl, wo = oracle.canonical_lloyd(x, c0, 3, 0.0)[:3]
            assert np.array_equal(l_out[i], wl), i
            assert np.array_equal(c_out[i].view(np.uint8), wc.view(np.uint8)), i
            assert obj[i] == wo, i


def test_batch_matrix_core_tile_exact_on_ties(engine, oracle):
    """The certified matrix-core variant on integer data with duplicate centroids
    (exact ties everywhere: candidate re-score and full re-scan paths) equals the
    canonical result; instances it does not apply to (k > 128) take a VALU tile."""
    from onedal_amd.engine import KM_OPT_BATCH_TILE

    rng = np.random.default_rng(41)
    inst = []
    for i, (n, d, k) in enumerate([(3000, 3, 5), (5000, 16, 40), (4000, 64, 128), (2500, 100, 77),
                                   (1500, 128, 128), (2000, 7, 129), (1000, 32, 1)]):
        x = rng.integers(-3, 4, size=(n, d)).astype(np.float32)
        idx = rng.choice(n, size=max(1, k // 2), replace=False)
        c0 = x[np.resize(idx, k)].copy()  # duplicated centroids
        inst.append((x, c0))
    try:
        engine.set_option(KM_OPT_BATCH_TILE, 8)  # variant 7
        c_out, l_out, obj, its = run_batch(engine, inst, 3, graph=False)
    finally:
        engine.set_option(KM_OPT_BATCH_TILE, 0)
    for i, (x, c0) in enumerate(inst):
        wc, wl, wo = oracle.canonical_lloyd(x, c0, 3, 0.0)[:3]
        assert np.array_equal(l_out[i], wl), i
        assert np.array_equal(c_out[i].view(np.uint8), wc.view(np.uint8)), i
        assert obj[i] == wo, i


@pytest.mark.parametrize("scale", [1e-12, 1.0, 1e17], ids=["tiny", "unit", "beyond_norm_range"])
def test_batch_matrix_core_tile_magnitudes(engine, oracle, scale):
    """The certified matrix-core variant at small, unit and large magnitudes (rows
    and centroids around 1e-12, 1 and 1e17 per feature, so squared norms from 1e-22
    to 1e38) equals the canonical result."""
    from onedal_amd.engine import KM_OPT_BATCH_TILE

    rng = np.random.default_rng(43)
    inst = []
    for i, (n, d, k) in enumerate([(3000, 16, 20), (2000, 64, 50), (2500, 128, 100), (1500, 32, 7),
                                   (2000, 100, 64)]):
        x = (gaussian_mixture(n, d, max(1, k // 2), seed=4300 + i, center_range=10.0,
                              sigma=(0.5, 2.0)) * np.float32(scale)).astype(np.float32)
        c0 = x[rng.choice(n, size=k, replace=False)].copy()
        inst.append((x, c0))
    try:
        engine.set_option(KM_OPT_BATCH_TILE, 8)  # variant 7
        c_out, l_out, obj, its = run_batch(engine, inst, 3, graph=False)
    finally:
        engine.set_option(KM_OPT_BATCH_TILE, 0)
    for i, (x, c0) in enumerate(inst):
        wc, wl, wo = oracle.canonical_lloyd(x, c0, 3, 0.0)[:3]
        assert np.array_equal(l_out[i], wl), i
        assert np.array_equal(c_out[i].view(np.uint8), wc.view(np.uint8)), i
        assert obj[i] == wo, i
