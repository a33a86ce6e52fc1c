"""The CPU baseline that bench.py times (oracle/kmeans_cpu_fast.cpp: the
reference's CPU Lloyd algorithm with a vectorised X.C^T kernel) computes what the
reference's algorithm computes: on the reference's gold dataset it reaches the
gold centroids, responses and objective at the reference's tolerances
(algo/kmeans/test/fixture.hpp:156-203), and on a Gaussian mixture it matches the
DAAL restatement (oracle.daal_lloyd) to float rounding, labels equal outside the
float band."""

import numpy as np

from conftest import gaussian_mixture
from test_oracle_golden import load_gold, match_map, rel_close


def test_cpu_baseline_gold_dataset(oracle):
    g, x, c0, ec, er = load_gold(np.float32)
    c, lab, obj = oracle.cpu_fast_lloyd(x, c0, g["max_iteration_count"])
    mm = match_map(ec, c)
    assert rel_close(ec[mm], c, 1e-5)
    assert np.array_equal(er, mm[lab])
    assert abs(obj - g["expected_objective"]) / g["expected_objective"] < 1e-5


def test_cpu_baseline_matches_daal_restatement(oracle):
    for (n, d, k, seed) in [(5000, 20, 10, 0), (3000, 64, 100, 1), (2000, 33, 37, 2)]:
        x = gaussian_mixture(n, d, k, seed=seed, center_range=10.0)
        c0 = x[:k].copy()
        c, lab, obj = oracle.cpu_fast_lloyd(x, c0, 3, nthreads=3, block_rows=96)
        dc, dl, do, _ = oracle.daal_lloyd(x, c0, 3, 0.0, stop_rule=2, block_size=512)
        _, b1, b2 = oracle.fp64_assign(x, dc)
        x64 = x.astype(np.float64)
        xn = (x64 * x64).sum(1)
        cl = oracle.fp64_assign(x, dc)[0]
        cn = (dc.astype(np.float64) ** 2).sum(1)[cl]
        band = (b2 - b1) < 4.0 * d * 2.0 ** -24 * (xn + cn)
        assert np.array_equal(lab[~band], dl[~band])
        assert np.allclose(c, dc, rtol=1e-5, atol=1e-5)
        assert abs(obj - do) / do < 1e-4
    assert oracle.cpu_fast_isa() in ("avx512", "avx2", "portable")


def test_cpu_baseline_without_final_pass(oracle):
    x = gaussian_mixture(4000, 16, 8, seed=5)
    c0 = x[:8].copy()
    c1, l1, o1 = oracle.cpu_fast_lloyd(x, c0, 2, final=False)
    c2, l2, o2 = oracle.cpu_fast_lloyd(x, c0, 2, final=True)
    assert o1 == 0.0 and o2 > 0.0
    assert np.array_equal(c1, c2)
