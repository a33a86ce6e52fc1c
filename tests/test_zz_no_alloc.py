"""km_lloyd_infer_f32_noalloc against km_lloyd_infer_f32 (same device data).

Bar: labels and objective bit-identical to the allocating entry point on every
assignment path (generic tile kernel, FP32 matrix cores, BF16 matrix cores with
exact re-score), over two calls in a row with the second on a smaller input, on
inputs that take one row block and several (rows padded to a multiple of four
floats, too); one small case also against the oracle's exact assignment.
Refusals (a k x d that leaves no room in the fixed scratch buffer, host
pointers) are checked before any kernel runs."""

import ctypes

import numpy as np
import pytest

from conftest import gaussian_mixture

pytestmark = pytest.mark.gpu

PATH_OPT = {"generic": 1, "mfma_f32": 2, "bf16": 0}


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _compare(engine, x, c, bf16, tag):
    import torch

    xd, cd = _dev(x), _dev(c)
    want_lab, want_obj = engine.infer_device(xd, cd, objective=True, bf16=bf16)
    want_path = engine.stats()["assign_path"]
    labels = torch.full((x.shape[0],), -7, dtype=torch.int32, device="cuda:0")
    obj = engine.infer_device_noalloc(xd, cd, labels, objective=True, bf16=bf16)
    assert engine.stats()["assign_path"] == want_path, tag
    got = labels.cpu().numpy()
    want = want_lab.cpu().numpy()
    mism = np.flatnonzero(got != want)
    assert mism.size == 0, "%s: label mismatch at rows %s" % (tag, mism[:10])
    assert obj == want_obj, (tag, obj, want_obj)
    # labels only (no objective requested) are the same too
    labels.fill_(-7)
    assert engine.infer_device_noalloc(xd, cd, labels, objective=False, bf16=bf16) is None
    assert np.array_equal(labels.cpu().numpy(), want), tag
    return got, obj


# (path, first call (n, d, k), second call (n, d, k)); the second is smaller.
# mfma_f32 at n=1.5e6, d=21 pads rows to 24 floats and runs in several blocks;
# bf16 at n=3e5, d=128 runs in several blocks.
CASES = [
    ("mfma_f32", (1_500_000, 21, 64), (70_000, 21, 64)),
    ("mfma_f32", (300_000, 64, 256), (50_001, 64, 100)),
    ("bf16", (300_000, 128, 300), (40_000, 128, 300)),
    ("bf16", (120_000, 48, 1000), (9_000, 48, 17)),
    ("generic", (100_000, 300, 64), (5_000, 300, 40)),
]


@pytest.mark.parametrize("path,first,second", CASES,
                         ids=["%s-n%d_d%d_k%d" % (p, *f) for p, f, _ in CASES])
def test_noalloc_equals_infer(engine, path, first, second):
    from onedal_amd.engine import KM_OPT_ASSIGN_PATH

    engine.set_option(KM_OPT_ASSIGN_PATH, PATH_OPT[path])
    bf16 = path == "bf16"
    try:
        for call, (n, d, k) in enumerate((first, second)):
            x = gaussian_mixture(n, d, k, seed=100 + call, center_range=20.0, sigma=(0.5, 2.0))
            rng = np.random.default_rng(7 + call)
            c = x[rng.choice(n, k, replace=False)] + 0.01
            tag = "%s call %d n=%d d=%d k=%d" % (path, call, n, d, k)
            print(tag, flush=True)
            _compare(engine, x, c.astype(np.float32), bf16, tag)
    finally:
        engine.set_option(KM_OPT_ASSIGN_PATH, 0)


def test_noalloc_vs_oracle(engine, oracle):
    n, d, k = 20_000, 32, 40
    x = gaussian_mixture(n, d, k, seed=5)
    c = x[:k] + 0.5
    got, obj = _compare(engine, x, c, False, "oracle case")
    want, _ = oracle.canonical_assign(x, c)
    assert np.array_equal(got, want)
    assert obj == oracle.canonical_objective(x, c, got)


def test_noalloc_refusals(engine):
    import torch

    from onedal_amd.engine import KM_FLAG_DEVICE_POINTERS, KM_INVALID_ARGUMENT, load_library

    L = load_library()
    x = torch.zeros((8, 256), dtype=torch.float32, device="cuda:0")
    lab = torch.zeros(8, dtype=torch.int32, device="cuda:0")
    obj = ctypes.c_double()
    # centroid-side workspace of k x d = 65536 x 256 floats fills the whole buffer:
    # refused before any kernel runs (the centroid pointer is never read)
    st = L.km_lloyd_infer_f32_noalloc(engine.ctx, x.data_ptr(), 8, 256, x.data_ptr(), 65536,
                                      KM_FLAG_DEVICE_POINTERS, lab.data_ptr(), ctypes.byref(obj))
    assert st == KM_INVALID_ARGUMENT
    assert b"too large" in L.km_last_error(engine.ctx)
    # host pointers are refused
    st = L.km_lloyd_infer_f32_noalloc(engine.ctx, x.data_ptr(), 8, 256, x.data_ptr(), 4, 0,
                                      lab.data_ptr(), ctypes.byref(obj))
    assert st == KM_INVALID_ARGUMENT
    assert b"device pointers" in L.km_last_error(engine.ctx)
    # the context still works afterwards
    xs = gaussian_mixture(5000, 16, 8, seed=3)
    _compare(engine, xs, xs[:8].copy(), False, "after refusals")
