"""The plain FP32 workload's own recipe on the GPU, against the oracle.

Recipe (the drop-in descriptor workload, scaled from n=1e7 to n=2e5 so the CPU
restatements finish in seconds): 256 Gaussians, centres U[-20,20]^64, per-cluster
sigma U[0.5,2], k=256 initial centroids drawn as distinct random rows, exactly 20
Lloyd iterations, accuracy threshold 0.

* Bit-exact against the engine's contract restated on the CPU (canonical_lloyd).
* Against the reference's CPU algorithm (the DAAL restatement), in lockstep: at
  every one of the 20 iterates the engine's labels equal the DAAL labels except
  rows whose FP64 top-2 gap is below eps * (|x|^2 + |c_1|^2), eps = 4 d 2^-24
  (both sides round the same distances differently in float), and the centroid
  update of every cluster no such row touches agrees to 1e-5.  The reference pins
  labels the same way: equal to the brute force away from ties
  (backend/gpu/test/cluster_updater_dpc.cpp:537-557).
* Free-running, the two 20-iteration trajectories part at the first in-band row
  whose label differs (a point on a bisector moves a centroid by |x - c| / n_j),
  so the end-to-end comparison states the mismatch count and requires the
  objectives to agree to 1e-4.
"""

import numpy as np
import pytest

from conftest import gaussian_mixture
from test_gpu_parity import band_mask

pytestmark = pytest.mark.gpu

N, D, K, ITERS = 200_000, 64, 256, 20


@pytest.fixture(scope="module")
def fp32_workload():
    x = gaussian_mixture(N, D, K, seed=2, center_range=20.0, sigma=(0.5, 2.0))
    rng = np.random.default_rng(2)
    c0 = x[rng.choice(N, size=K, replace=False)].copy()
    return x, c0


def exact_objective(x, c, labels):
    diff = x.astype(np.float64) - c.astype(np.float64)[labels]
    return float(np.einsum("ij,ij->", diff, diff))


def test_fp32_workload_bit_exact_vs_canonical(engine, oracle, fp32_workload):
    from test_gpu_parity import assert_bit_exact, set_path

    x, c0 = fp32_workload
    set_path(engine, "mfma_f32")
    got = engine.train(x, c0, ITERS, 0.0)
    assert engine.stats()["assign_path"] == 2
    want = oracle.canonical_lloyd(x, c0, ITERS, 0.0)
    assert_bit_exact(got, want)
    # the per-iteration objectives of the oracle end at the returned objective's run
    assert len(want[4]) == ITERS


def test_fp32_workload_lockstep_vs_daal(engine, oracle, fp32_workload):
    from test_gpu_parity import set_path

    x, c0 = fp32_workload
    set_path(engine, "mfma_f32")
    c = c0
    in_band_total = 0
    refills = 0
    for it in range(ITERS):
        lab, obj = engine.infer(x, c, objective=True)
        dc, dl, dobj, _ = oracle.daal_lloyd(x, c, 0, 0.0, stop_rule=2, block_size=512)
        band = band_mask(oracle, x, c)
        mism = lab != dl
        outside = np.flatnonzero(mism & ~band)
        assert outside.size == 0, (it, outside[:10])
        in_band_total += int(mism.sum())
        # the engine's objective is the FP64 sum of exact-form distances; DAAL's
        # adds float distances in float block partials (about 1e-5 low here)
        assert abs(obj - exact_objective(x, c, lab)) / obj < 1e-12
        assert abs(obj - dobj) / dobj < 5e-5, (it, obj, dobj)
        # one Lloyd update from the same centroids: every cluster no differing row
        # touches agrees to 1e-5 (FP64 sums against DAAL's float per-thread sums);
        # clusters refilled with the farthest rows agree bit for bit
        c_next = engine.train(x, c, 1, 0.0)[0]
        dnext = oracle.daal_lloyd(x, c, 1, 0.0, stop_rule=2, block_size=512)[0]
        touched = np.zeros(K, bool)
        touched[lab[mism]] = True
        touched[dl[mism]] = True
        counts = np.bincount(lab, minlength=K)
        keep = ~touched & (counts > 0)
        scale = np.maximum(1.0, np.abs(dnext[keep]).max(axis=1, keepdims=True))
        err = np.abs(c_next[keep].astype(np.float64) - dnext[keep]) / scale
        assert err.max() < 1e-5, (it, err.max())
        refilled = np.flatnonzero(counts == 0)
        assert np.array_equal(c_next[refilled], dnext[refilled]), it
        refills += refilled.size
        c = c_next
    print("lockstep: %d iterations, %d label differences in total, all inside the float band; "
          "%d clusters refilled" % (ITERS, in_band_total, refills))


def test_fp32_workload_free_running_vs_daal(engine, oracle, fp32_workload):
    from test_gpu_parity import set_path

    x, c0 = fp32_workload
    set_path(engine, "mfma_f32")
    c, lab, obj, it = engine.train(x, c0, ITERS, 0.0)
    dc, dl, dobj, dit = oracle.daal_lloyd(x, c0, ITERS, 0.0, stop_rule=2, block_size=512)
    assert it == dit == ITERS
    rel = abs(obj - dobj) / dobj
    print("free-running: %d of %d labels differ, objective rel diff %.3e" % (int((lab != dl).sum()), N, rel))
    assert rel < 1e-4
