"""GPU parity tests: the HIP engine (through the C ABI) against the oracle.

Bar: bit-exact against the canonical restatement (labels, centroids,
objective, iteration count) on every path; against the reference's golden
vectors at the reference's own tolerances; against the DAAL restatement to
rounding."""

import numpy as np
import pytest

from conftest import gaussian_mixture
from test_oracle_golden import load_gold, match_map, rel_close


def band_mask(oracle, x, c):
    """Rows whose FP64 top-2 gap against centroids c is below
    eps * (|x|^2 + |c_1|^2), eps = 4 d u (u = 2^-24 for float, 2^-53 for double):
    the rows whose label two correct F-precision evaluations may order differently."""
    lab, b1, b2 = oracle.fp64_assign(x, c)
    x64 = x.astype(np.float64)
    c64 = c.astype(np.float64)
    xn = np.einsum("ij,ij->i", x64, x64)
    cn = np.einsum("ij,ij->i", c64, c64)
    u = 2.0 ** -24 if x.dtype == np.float32 else 2.0 ** -53
    return (b2 - b1) < 4.0 * x.shape[1] * u * (xn + cn[lab])

pytestmark = pytest.mark.gpu

PATHS = {"generic": 1, "mfma_f32": 2, "bf16": 0}


def set_path(engine, path):
    """Select the assignment path; returns the bf16 flag to pass to train/infer."""
    from onedal_amd.engine import KM_OPT_ASSIGN_PATH

    engine.set_option(KM_OPT_ASSIGN_PATH, PATHS[path])
    return path == "bf16"


def assert_bit_exact(got, want):
    c, lab, obj, it = got
    wc, wl, wo, wi = want[:4]
    assert it == wi
    mism = np.flatnonzero(lab != wl)
    assert mism.size == 0, "label mismatch at rows %s" % mism[:10]
    assert np.array_equal(c.view(np.uint8), wc.view(np.uint8)), np.max(np.abs(c - wc))
    assert obj == wo, (obj, wo)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("path", ["generic", "mfma_f32", "bf16"])
def test_gold_dataset(engine, oracle, dtype, path):
    bf16 = set_path(engine, path)
    g, x, c0, ec, er = load_gold(dtype)
    res = engine.train(x, c0, g["max_iteration_count"], 0.0, bf16=bf16)
    c, lab, obj, it = res
    # reference tolerances (test/fixture.hpp:156-203)
    mm = match_map(ec, c)
    assert rel_close(ec[mm], c, 1e-7)
    assert np.array_equal(er, mm[lab])
    assert abs(obj - g["expected_objective"]) / g["expected_objective"] < 1e-7
    # bit-exact against the canonical restatement
    assert_bit_exact(res, oracle.canonical_lloyd(x, c0, g["max_iteration_count"], 0.0))
    # inference with the trained model (test/fixture.hpp:185-193)
    ilab, iobj = engine.infer(x, c, objective=True, bf16=bf16)
    assert np.array_equal(er, mm[ilab])
    assert iobj == obj


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("path", ["generic", "mfma_f32", "bf16"])
def test_empty_clusters_fixture(engine, oracle, dtype, path):
    bf16 = set_path(engine, path)
    x = np.array([-10, -9.5, -9, -8.5, -8, -1, 1, 9, 9.5, 10], dtype).reshape(10, 1)
    c0 = np.array([-10, -10, -10], dtype).reshape(3, 1)
    c, lab, obj, it = engine.train(x, c0, 1, 0.0, bf16=bf16)
    assert rel_close(c.ravel(), np.array([-1.65, 10, 9.5], dtype), 5 * np.finfo(dtype).eps * 100)
    assert lab.tolist() == [0, 0, 0, 0, 0, 0, 0, 2, 2, 1]
    assert it == 1
    assert engine.stats()["empty_cluster_events"] == 1
    assert_bit_exact((c, lab, obj, it), oracle.canonical_lloyd(x, c0, 1, 0.0))


@pytest.mark.parametrize("path", ["generic", "mfma_f32", "bf16"])
@pytest.mark.parametrize("thr", [0.0, 1e-12], ids=["no_stop_rule", "stop_rule"])
def test_empty_clusters_many_with_ties(engine, oracle, path, thr):
    """Many empty clusters at once, with tied d64 values (integer rows with many
    duplicates; initial centroids repeated, so the repeats start empty) over rows
    spread across many workgroups: the device-side selection (radix select,
    in-order collection, pick order by d64 then row) equals the canonical
    restatement, with and without the stop rule's objective."""
    bf16 = set_path(engine, path)
    rng = np.random.default_rng(77)
    n, d, k = 60000, 4, 96
    x = rng.integers(-6, 7, size=(n, d)).astype(np.float32)
    idx = rng.choice(n, size=k // 3, replace=False)
    c0 = x[np.resize(idx, k)].copy()  # every centroid three times
    got = engine.train(x, c0, 4, thr, bf16=bf16)
    assert engine.stats()["empty_cluster_events"] >= 1
    assert_bit_exact(got, oracle.canonical_lloyd(x, c0, 4, thr))


SHAPES = [
    # n, d, k, iters
    (1000, 1, 3, 5),
    (3000, 3, 10, 5),
    (10000, 20, 10, 8),
    (8000, 33, 37, 5),
    (12000, 64, 64, 5),
    (6000, 100, 17, 4),
    (6000, 128, 130, 3),
    (3000, 200, 50, 3),
    (2000, 256, 257, 2),
    (2000, 300, 40, 2),
    (700, 7, 700, 2),
]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("path", ["generic", "mfma_f32", "bf16"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "n%d_d%d_k%d" % s[:3])
def test_train_bit_exact_vs_canonical(engine, oracle, dtype, path, shape):
    n, d, k, iters = shape
    bf16 = set_path(engine, path)
    x = gaussian_mixture(n, d, k, seed=n + d + k, center_range=10.0, sigma=(0.5, 2.0), dtype=dtype)
    rng = np.random.default_rng(k)
    c0 = x[rng.choice(n, size=k, replace=False)].copy()
    got = engine.train(x, c0, iters, 0.0, bf16=bf16)
    assert_bit_exact(got, oracle.canonical_lloyd(x, c0, iters, 0.0))


BF16_EDGE_CASES = [
    # name, n, d, k, iters, spread: rows nearly identical (spread 1e-3) put every
    # centroid inside every row's BF16 error band (collect and full re-score
    # paths); k = 1500 spans 48 centroid blocks; k = 130 leaves padding blocks
    ("all_centroids_in_band", 6000, 64, 300, 3, 1e-3),
    ("k1500_48_blocks", 30000, 32, 1500, 3, None),
    ("k_not_multiple_of_128", 20000, 128, 130, 4, None),
]


@pytest.mark.parametrize("case", BF16_EDGE_CASES, ids=lambda c: c[0])
def test_bf16_edge_cases_bit_exact(engine, oracle, case):
    """BF16 matrix-core distances with exact re-score at the band's extremes: the
    result stays the canonical one."""
    _, n, d, k, iters, spread = case
    bf16 = set_path(engine, "bf16")
    if spread is None:
        x = gaussian_mixture(n, d, k, seed=n + k, center_range=10.0, sigma=(0.5, 2.0))
    else:
        rng = np.random.default_rng(3)
        x = (1.0 + spread * rng.standard_normal((n, d))).astype(np.float32)
    rng = np.random.default_rng(k + 1)
    c0 = x[rng.choice(n, size=k, replace=False)].copy()
    got = engine.train(x, c0, iters, 0.0, bf16=bf16)
    assert_bit_exact(got, oracle.canonical_lloyd(x, c0, iters, 0.0))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_train_vs_daal_restatement(engine, oracle, dtype):
    """The reference's CPU workload shape (n=1e4, d=20, k=10, first-k init,
    threshold 0): the engine against the reference's CPU algorithm, to rounding."""
    x = gaussian_mixture(10000, 20, 10, seed=0, center_range=10.0, dtype=dtype)
    c0 = x[:10].copy()
    c, lab, obj, it = engine.train(x, c0, 50, 0.0)
    dc, dl, do, di = oracle.daal_lloyd(x, c0, 50, 0.0, stop_rule=2, block_size=512)
    assert it == 50
    # labels equal except rows whose FP64 top-2 gap is inside the band
    # eps * (|x|^2 + |c_1|^2), eps = 4 d 2^-24 (F = float; 4 d 2^-53 for double)
    band = band_mask(oracle, x, dc)
    assert band.mean() < 1e-3  # almost every row is compared
    assert np.array_equal(lab[~band], dl[~band])
    print("daal restatement: %d labels differ, %d rows inside the band" % (int((lab != dl).sum()),
                                                                          int(band.sum())))
    assert np.allclose(c, dc, rtol=1e-5, atol=1e-5)
    assert abs(obj - do) / do < 1e-5


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("path", ["generic", "mfma_f32", "bf16"])
def test_infer_bit_exact(engine, oracle, dtype, path):
    bf16 = set_path(engine, path)
    x = gaussian_mixture(20000, 48, 33, seed=11, dtype=dtype)
    c = gaussian_mixture(33, 48, 33, seed=12, dtype=dtype)
    lab, obj = engine.infer(x, c, objective=True, bf16=bf16)
    wl, _ = oracle.canonical_assign(x, c)
    assert np.array_equal(lab, wl)
    assert obj == oracle.canonical_objective(x, c, wl)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_duplicate_centroids_tie_to_lowest_index(engine, oracle, dtype):
    """Exact ties (duplicate centroids) must go to the lowest index, as in
    select_min_distance (kernels_fp_impl.hpp:198-202); exercises the re-score
    kernels behind the FP32 and the FP64 matrix-core assignment."""
    set_path(engine, "mfma_f32")
    x = gaussian_mixture(5000, 16, 4, seed=5, dtype=dtype)
    c = np.concatenate([x[:4], x[:4]]).copy()  # centroid j+4 duplicates j
    lab, _ = engine.infer(x, c)
    assert lab.max() < 4
    assert np.array_equal(lab, oracle.canonical_assign(x, c)[0])
    st = engine.stats()
    # every row has two candidates inside the band: none is settled without a re-score
    assert st["refined_full_rows"] + st["refined_candidate_rows"] == 5000


def test_stress_identity_large(engine):
    """test/fixture.hpp:289-308 at full size: n=8192, d=2048, k=8192, C0 = X."""
    rng = np.random.default_rng(7777)
    x = rng.uniform(-0.2, 0.5, size=(8192, 2048)).astype(np.float32)
    c, lab, obj, it = engine.train(x, x, 1, 0.0)
    assert np.array_equal(lab, np.arange(8192))
    assert np.array_equal(c, x)
    assert obj == 0.0


def test_block_one_cluster_mean(engine):
    """test/fixture.hpp:254-287 at full size: n=1M, d=256, k=1 -> column mean."""
    rng = np.random.default_rng(7777)
    x = rng.uniform(-0.2, 0.5, size=(1 << 20, 256)).astype(np.float32)
    c, lab, obj, it = engine.train(x, x[:1].copy(), 1, 0.0)
    mean = x.astype(np.float64).mean(0)
    assert np.max(np.abs(c[0] - mean) / np.abs(mean)) < 1e-6
    ref_obj = ((x.astype(np.float64) - c[0].astype(np.float64)) ** 2).sum()
    assert abs(obj - ref_obj) / ref_obj < 1e-9
    assert not lab.any()


def test_repeatable_bitwise(engine):
    x = gaussian_mixture(50000, 64, 100, seed=9)
    c0 = x[:100].copy()
    a = engine.train(x, c0, 5, 0.0)
    b = engine.train(x, c0, 5, 0.0)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_stop_rule_gpu_objective(engine, oracle):
    """accuracy_threshold > 0: stop when obj + thr > previous obj
    (train_kernel_lloyd_dense_dpc.cpp:162-166)."""
    x = gaussian_mixture(4000, 5, 6, seed=21)
    c0 = x[:6].copy()
    got = engine.train(x, c0, 100, 1e-3)
    want = oracle.canonical_lloyd(x, c0, 100, 1e-3)
    assert got[3] < 100
    assert_bit_exact(got, want)


def test_badargs(engine):
    """test/badarg.cpp:114-218 through the C ABI."""
    from onedal_amd.engine import DomainError, InvalidArgument, Unimplemented

    x = np.zeros((10, 2), np.float32)
    c0 = np.zeros((3, 2), np.float32)
    with pytest.raises(DomainError):
        engine.train(x, np.zeros((0, 2), np.float32), 1, 0.0)  # k = 0
    with pytest.raises(DomainError):
        engine.train(x, c0, -1, 0.0)
    with pytest.raises(DomainError):
        engine.train(x, c0, 1, -0.5)
    with pytest.raises(InvalidArgument):
        engine.train(x[:2], c0, 1, 0.0)  # k > n
    with pytest.raises(DomainError):
        engine.train(np.zeros((0, 2), np.float32), c0, 1, 0.0)
    # still usable after errors
    c, lab, obj, it = engine.train(np.arange(20, dtype=np.float32).reshape(10, 2), c0 + [[0, 0], [5, 5], [15, 15]], 2, 0.0)
    assert it == 2


@pytest.mark.parametrize("shape", [(40000, 128, 1024, 3), (30000, 64, 256, 4), (20000, 32, 100, 4)],
                         ids=lambda s: "n%d_d%d_k%d" % s[:3])
def test_bf16_labels_identical_to_fp32(engine, oracle, shape):
    """The precision form's contract: BF16 distances + exact re-score give the same
    labels, centroids and objective as the FP32 path, bit for bit, on anisotropic
    Gaussian mixtures like the precision workload (centers U[-50,50]^d, per-dimension
    scale log-uniform in [0.1, 10])."""
    n, d, k, iters = shape
    x = gaussian_mixture(n, d, k, seed=d, center_range=50.0, anisotropic=True)
    rng = np.random.default_rng(1)
    c0 = x[rng.choice(n, size=k, replace=False)].copy()
    set_path(engine, "mfma_f32")
    a = engine.train(x, c0, iters, 0.0)
    set_path(engine, "bf16")
    b = engine.train(x, c0, iters, 0.0, bf16=True)
    st = engine.stats()
    assert st["assign_path"] == 3
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[0], b[0])
    assert a[2] == b[2] and a[3] == b[3]
    # near-ties were actually re-scored on this data
    assert st["refined_candidate_rows"] > 0
    assert_bit_exact(b, oracle.canonical_lloyd(x, c0, iters, 0.0))
