"""Serving on the GPU: km_batch_step_f32 (continuous-batch steps) and the km_server
request loop.  Every job's result must equal the single-instance engine on that
job alone, bit for bit (labels, centroid bytes, objective), whatever else is in
the mix, on either lane, from host or device memory."""

import threading
import time

import numpy as np
import pytest

from conftest import gaussian_mixture

pytestmark = pytest.mark.gpu

WAIT_MS = 120000


def _job_data(rng, seed, n, d, k, repeat_c0=False):
    x = gaussian_mixture(n, d, max(1, k // 2 + 1), seed=seed, center_range=10.0, sigma=(0.5, 2.0))
    if repeat_c0:
        idx = rng.choice(n, size=max(1, k // 2), replace=False)
        c0 = x[np.resize(idx, k)].copy()  # clusters that start empty
    else:
        c0 = x[rng.choice(n, size=k, replace=False)].copy()
    return x, c0


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


def test_batch_step_schedule_matches_train(engine):
    """Instances join, iterate by uneven chunks and finish at different steps (an
    instance is in both sets of its last step); infer instances finish in the
    step they join.  Each result equals km_lloyd_train_f32 / _infer_f32 alone."""
    import torch

    rng = np.random.default_rng(7)
    shapes = [(3000, 5, 7, 4), (5000, 16, 33, 6), (800, 1, 3, 2), (4100, 33, 64, 0),
              (2500, 128, 20, 5), (1200, 8, 1, 3), (6000, 17, 128, 7)]
    fits = []
    for i, (n, d, k, T) in enumerate(shapes):
        x, c0 = _job_data(rng, 100 + i, n, d, k, repeat_c0=(i % 3 == 0))
        fits.append(dict(x=x, c0=c0, T=T, join=i % 3, done=0,
                         X=torch.from_numpy(x).cuda(), C0=torch.from_numpy(c0).cuda(),
                         C=torch.empty((k, d), dtype=torch.float32, device="cuda"),
                         L=torch.empty(n, dtype=torch.int32, device="cuda"), finished=False))
    infers = []
    for i, (n, d, k) in enumerate([(2000, 12, 9), (700, 3, 800), (3000, 64, 40)]):
        x, _ = _job_data(rng, 200 + i, n, d, 4)
        c = rng.standard_normal((k, d)).astype(np.float32) * 5
        infers.append(dict(x=x, c=c, join=i, X=torch.from_numpy(x).cuda(), Cm=torch.from_numpy(c).cuda(),
                           L=torch.empty(n, dtype=torch.int32, device="cuda")))
    torch.cuda.synchronize()
    objs = {}
    for step in range(12):
        active = [f for f in fits if f["join"] <= step and not f["finished"]]
        iterating = [f for f in active if f["done"] < f["T"]]
        it_ids = {id(f) for f in iterating}
        chunk = 1 + step % 2
        if iterating:
            chunk = min([chunk] + [f["T"] - f["done"] for f in iterating])
        it_set = [(f["X"], f["C0"] if f["done"] == 0 else f["C"], f["C"]) for f in iterating]
        fin = []
        for f in active:
            after = f["done"] + (chunk if id(f) in it_ids else 0)
            if after == f["T"]:
                fin.append(("fit", f, (f["X"], f["C0"] if after == 0 else f["C"], f["C"], f["L"])))
        for q in infers:
            if q["join"] == step:
                fin.append(("infer", q, (q["X"], q["Cm"], None, q["L"])))
        if not it_set and not fin:
            continue
        o = engine.batch_step_device(it_set, chunk, [t[2] for t in fin])
        for f in iterating:
            f["done"] += chunk
        for (kind, rec, _), ob in zip(fin, o):
            objs[id(rec)] = ob
            if kind == "fit":
                rec["finished"] = True
    torch.cuda.synchronize()
    for f in fits:
        assert f["finished"]
        c, lab, ob, it = engine.train(f["x"], f["c0"], f["T"], 0.0)
        assert np.array_equal(lab, f["L"].cpu().numpy())
        assert _same(c, f["C"].cpu().numpy())
        assert ob == objs[id(f)]
    for q in infers:
        lab, ob = engine.infer(q["x"], q["c"], objective=True)
        assert np.array_equal(lab, q["L"].cpu().numpy())
        assert ob == objs[id(q)]


def test_batch_step_badargs(engine):
    import torch

    from onedal_amd.engine import DomainError, InvalidArgument

    x = torch.zeros((10, 3), dtype=torch.float32, device="cuda")
    c = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
    co = torch.empty_like(c)
    lab = torch.empty(10, dtype=torch.int32, device="cuda")
    with pytest.raises(DomainError):
        engine.batch_step_device([], 1, [])
    with pytest.raises(InvalidArgument):
        engine.batch_step_device([(x[:3], c, co)], 1, [])  # k > n
    with pytest.raises(DomainError):
        engine.batch_step_device([(x, c, co)], -1, [])
    # usable afterwards
    o = engine.batch_step_device([(x, c, co)], 1, [(x, co, None, lab)])
    assert len(o) == 1


def _make_mix(seed):
    """Fit and infer jobs of uneven sizes; the few largest take the large-job lane
    with the server configured below."""
    rng = np.random.default_rng(seed)
    dims = [1, 2, 5, 8, 13, 16, 32, 33, 64, 100, 128]
    jobs = []
    for i in range(30):
        d = dims[i % len(dims)]
        n = int(rng.integers(300, 12000))
        k = min(n, int(rng.integers(1, 100)))
        x, c0 = _job_data(rng, seed * 1000 + i, n, d, k, repeat_c0=(i % 6 == 0))
        if i % 4 == 3:
            c = rng.standard_normal((k, d)).astype(np.float32) * 4
            jobs.append(dict(kind="infer", x=x, c=c))
        else:
            jobs.append(dict(kind="fit", x=x, c0=c0, iters=int(rng.integers(0, 9))))
    # large-lane jobs: n*k*d >= LARGE_WORK
    for i, (n, d, k) in enumerate([(40000, 64, 96), (30000, 128, 70), (50000, 40, 130)]):
        x, c0 = _job_data(rng, seed * 1000 + 500 + i, n, d, k)
        if i == 2:
            jobs.append(dict(kind="infer", x=x, c=c0 + 0.5))
        else:
            jobs.append(dict(kind="fit", x=x, c0=c0, iters=5))
    for i, j in enumerate(jobs):
        j["device"] = (i % 2 == 1)
        j["bf16"] = (i % 3 == 0)
    return jobs


LARGE_WORK = 1.5e8


def _expected(engine, oracle, j):
    """The job's result from the oracle (the engine's contract restated on the
    CPU), checked first against the single-instance engine on the job alone."""
    if j["kind"] == "fit":
        c, lab, ob, it, _ = oracle.canonical_lloyd(j["x"], j["c0"], j["iters"], 0.0)
        got = engine.train(j["x"], j["c0"], j["iters"], 0.0)
        assert _same(got[0], c) and np.array_equal(got[1], lab) and got[2] == ob and got[3] == it
        return c, lab, ob, it
    lab, _ = oracle.canonical_assign(j["x"], j["c"])
    ob = oracle.canonical_objective(j["x"], j["c"], lab)
    got = engine.infer(j["x"], j["c"], objective=True)
    assert np.array_equal(got[0], lab) and got[1] == ob
    return None, lab, ob, 0


def _submit(server, j):
    import torch

    if j["device"]:
        x = torch.from_numpy(j["x"]).cuda()
        c = torch.from_numpy(j["c0"] if j["kind"] == "fit" else j["c"]).cuda()
        torch.cuda.synchronize()
    else:
        x, c = j["x"], (j["c0"] if j["kind"] == "fit" else j["c"])
    if j["kind"] == "fit":
        return server.submit_fit(x, c, j["iters"], bf16=j["bf16"])
    return server.submit_infer(x, c, bf16=j["bf16"])


def _host(a):
    if a is None:
        return None
    return a.cpu().numpy() if hasattr(a, "cpu") else a


def _check(j, res, want, large_work=None):
    c, lab, ob, it = want
    assert np.array_equal(_host(res["labels"]), lab)
    assert res["objective"] == ob
    assert res["iterations"] == it
    if c is not None:
        assert _same(_host(res["centroids"]), c)
    n, d = j["x"].shape
    k = (j["c0"] if j["kind"] == "fit" else j["c"]).shape[0]
    assert res["lane"] == (1 if n * k * d >= (large_work or LARGE_WORK) else 0)


def test_server_mix_equals_alone(engine, oracle):
    """Jobs submitted from three threads with staggered arrivals (the mix) and the
    same jobs one at a time (alone) give the oracle's results bit for bit."""
    from onedal_amd.engine import Server

    jobs = _make_mix(3)
    want = [_expected(engine, oracle, j) for j in jobs]
    ids = [None] * len(jobs)
    with Server(budget_bytes=8 << 30, large_job_work=int(LARGE_WORK)) as srv:
        errors = []

        def feed(part):
            try:
                for i in part:
                    ids[i] = _submit(srv, jobs[i])
                    time.sleep(0.002 * (i % 3))
            except Exception as e:  # pragma: no cover - reported below
                errors.append(e)

        threads = [threading.Thread(target=feed, args=(list(range(t, len(jobs), 3)),)) for t in range(3)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        assert not errors, errors
        results = [srv.wait(i, WAIT_MS) for i in ids]
        stats = srv.stats()
        for j, res, w in zip(jobs, results, want):
            assert res is not None
            _check(j, res, w)
        assert stats["completed"] == len(jobs) and stats["failed"] == 0
        assert stats["large_jobs"] == sum(1 for r in results if r["lane"] == 1) >= 3
        assert stats["max_step_jobs"] >= 2, stats  # jobs shared steps
        # alone: one job in the server at a time
        for i in range(0, len(jobs), 4):
            res = srv.wait(_submit(srv, jobs[i]), WAIT_MS)
            _check(jobs[i], res, want[i])
        assert srv.stats()["bytes_admitted"] == 0


def test_server_admission_budget(engine, oracle):
    """A budget of a few jobs' footprints: every job still completes with the
    single-instance result, the admitted footprint never exceeds the budget, and a
    job larger than the whole budget is refused at submission."""
    from onedal_amd.engine import EngineError, Server

    rng = np.random.default_rng(11)
    jobs = []
    for i in range(16):
        x, c0 = _job_data(rng, 900 + i, 20000, 16, 16)
        jobs.append(dict(kind="fit", x=x, c0=c0, iters=3, device=(i % 2 == 0), bf16=False))
    budget = 8 << 20
    with Server(budget_bytes=budget) as srv:
        ids = [_submit(srv, j) for j in jobs]
        results = [srv.wait(i, WAIT_MS) for i in ids]
        stats = srv.stats()
        assert stats["peak_bytes_admitted"] <= budget
        assert stats["completed"] == len(jobs)
        for j, res in zip(jobs, results):
            _check(j, res, _expected(engine, oracle, j))
        assert max(r["queue_ms"] for r in results) > 0.0
        big = gaussian_mixture(200000, 16, 4, seed=1)
        with pytest.raises(EngineError) as e:
            srv.submit_fit(big, big[:4].copy(), 2)
        assert e.value.status == 3  # KM_DEVICE_BAD_ALLOC
        # a pending wait with a zero timeout, then the result
        jid = _submit(srv, jobs[0])
        r = srv.wait(jid, 0)
        if r is None:
            r = srv.wait(jid, WAIT_MS)
        _check(jobs[0], r, _expected(engine, oracle, jobs[0]))


def test_server_badargs():
    from onedal_amd.engine import DomainError, InvalidArgument, Server

    x = np.zeros((10, 3), np.float32)
    c = np.zeros((4, 3), np.float32)
    with Server(budget_bytes=1 << 30) as srv:
        with pytest.raises(InvalidArgument):
            srv.submit_fit(x[:3], c, 2)  # k > n
        with pytest.raises(DomainError):
            srv.submit_fit(x, c, -1)
        with pytest.raises(InvalidArgument):
            srv.wait(12345, 0)
        jid = srv.submit_fit(x, c, 1)
        r = srv.wait(jid, WAIT_MS)
        assert r["iterations"] == 1
    with pytest.raises(DomainError):
        Server(budget_bytes=0)


def test_server_close_finishes_outstanding_jobs(engine):
    """Destroying the server runs every submitted job to completion first."""
    from onedal_amd.engine import Server

    rng = np.random.default_rng(5)
    srv = Server(budget_bytes=1 << 30)
    jobs = []
    for i in range(6):
        x, c0 = _job_data(rng, 40 + i, 5000, 8, 10)
        jid = srv.submit_fit(x, c0, 4)
        jobs.append((x, c0, srv._held[jid][1]))
    srv.close()
    for x, c0, (cout, labels) in jobs:
        c, lab, _, _ = engine.train(x, c0, 4, 0.0)
        assert np.array_equal(labels, lab) and _same(cout, c)


def test_server_serving_recipe_vs_oracle(engine, oracle):
    """The serving workload's recipe at reduced sizes: Poisson arrivals, 70% fit
    jobs of 20 iterations and 30% infer jobs, n log-uniform (here [1e3, 3e4]), d
    from {8, 16, 32, 64, 128}, k uniform in [4, 128], Gaussian-mixture data with k
    centres per job, BF16+refine requested, plus large jobs (here n=6e4, d=128,
    k=512) on the large-job lane; every result equals the oracle bit for bit."""
    from onedal_amd.engine import Server

    rng = np.random.default_rng(2024)
    jobs = []
    for i in range(40):
        n = int(np.exp(rng.uniform(np.log(1e3), np.log(3e4))))
        d = int(rng.choice([8, 16, 32, 64, 128]))
        k = int(rng.integers(4, 129))
        x = gaussian_mixture(n, d, k, seed=5000 + i, center_range=10.0)
        if rng.uniform() < 0.7:
            c0 = x[rng.choice(n, size=k, replace=False)].copy()
            jobs.append(dict(kind="fit", x=x, c0=c0, iters=20))
        else:
            c = gaussian_mixture(k, d, k, seed=6000 + i, center_range=10.0)
            jobs.append(dict(kind="infer", x=x, c=c))
    for i in range(2):
        x = gaussian_mixture(60000, 128, 512, seed=7000 + i, center_range=10.0)
        c0 = x[rng.choice(60000, size=512, replace=False)].copy()
        jobs.insert(10 + 15 * i, dict(kind="fit", x=x, c0=c0, iters=20))
    for i, j in enumerate(jobs):
        j["device"] = (i % 2 == 0)
        j["bf16"] = True
    want = [_expected(engine, oracle, j) for j in jobs]
    large = 1e9
    with Server(budget_bytes=4 << 30, large_job_work=int(large)) as srv:
        ids = []
        for j in jobs:
            ids.append(_submit(srv, j))
            time.sleep(rng.exponential(1.0 / 400.0))
        results = [srv.wait(i, WAIT_MS) for i in ids]
        stats = srv.stats()
    for j, res, w in zip(jobs, results, want):
        assert res is not None
        _check(j, res, w, large_work=large)
    assert stats["completed"] == len(jobs) and stats["failed"] == 0
    assert stats["large_jobs"] == 2
    assert stats["max_step_jobs"] >= 2, stats
