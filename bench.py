"""Benchmark of the MI355X k-means Lloyd engine.

One step = one full Lloyd training (fixed iteration count) over a resident
synthetic Gaussian-mixture matrix, through the C ABI with device pointers.
value = point-iterations/s = n * iterations * steps / elapsed.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload fp32_c2]

The batched workload (batched_c4) trains thousands of instances per call; the
serving workload (serving_c5) sweeps offered request rates through the km_server
request loop and reports throughput with p50/p99 job latency per rate.
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("k-means point-iters/s + %roofline, bf16+refine vs fp32 path; p50/p99 job latency")
PEAK = {"fp32_mfma_tflops": 157.3, "bf16_mfma_tflops": 2500.0, "hbm_gbs": 8000.0,
        "fp32_vector_tflops": 157.3}

# HBM bytes per launch of the main assignment kernel, from two rocprofv3 counter
# passes (FETCH_SIZE, WRITE_SIZE) scaled by a 1 GiB copy of the same run
# (profiles/traffic_precision_n20M_d128_k1024.md); hardware counters cannot be read
# from inside this process.
MEASURED_TRAFFIC = {"precision_c3": {"bytes": 5.538e9, "read": 5.300e9, "write": 0.238e9,
                                     "source": "profiles/traffic_precision_n20M_d128_k1024.md"}}

WORKLOADS = {
    # Precision form: n=2e7, d=128, k=1024, 1024 anisotropic Gaussians (per-dimension
    # scale log-uniform in [0.1,10], centers U[-50,50]^d), random-row init, 20 iterations,
    # BF16 matrix-core distances with exact FP32 re-score, FP64 centroid/objective sums
    "precision_c3": dict(n=20_000_000, d=128, k=1024, iters=20, center_range=50.0, sigma=None,
                         anisotropic=True, bf16=True),
    # Plain FP32 Lloyd (drop-in descriptor path): n=1e7, d=64, k=256, 256 Gaussians,
    # centers U[-20,20]^d, per-cluster sigma U[0.5,2], random-row init, 20 iterations
    "fp32_c2": dict(n=10_000_000, d=64, k=256, iters=20, center_range=20.0, sigma=(0.5, 2.0),
                    anisotropic=False, bf16=False),
}

# Batched form: independent instances, n log-uniform in [1e3, 1e6], d from
# {8, 16, 32, 64, 128}, k uniform in [4, 128], k Gaussians per instance (centers
# U[-10,10]^d, sigma 1), random distinct rows as initial centroids, 30 iterations,
# all instances in one batched call (each iteration replayed as a HIP graph).
BATCHED = {"batched_c4": dict(count=4096, n_lo=1e3, n_hi=1e6, dims=(8, 16, 32, 64, 128), k_lo=4,
                              k_hi=128, iters=30)}

# Serving form: a request loop with Poisson arrivals, swept over offered rates;
# 70% fit jobs (20 Lloyd iterations) and 30% infer jobs (labels and objective
# against resident centroids); job sizes as the batched form plus 5% large jobs
# (n=5e6, d=128, k=512); float32 data, BF16+refine distances requested (the
# large-job lane runs them; the batched lane's exact form gives the same labels);
# admission by a 200 GB HBM budget.  Datasets are generated once into HBM and
# each request names one of them.
SERVING = {"serving_c5": dict(rates=(50, 100, 200, 400, 1000, 2000), window_s=2.0, min_jobs=200,
                              max_jobs=4000, fit_frac=0.7, large_frac=0.05, iters=20, budget=200e9,
                              pool_small=96, pool_large=2, large=(5_000_000, 128, 512), n_lo=1e3,
                              n_hi=1e6, dims=(8, 16, 32, 64, 128), k_lo=4, k_hi=128)}

# Initial centroids: k-means++ (plus_plus_dense) on the FP32 Lloyd workload's data
# (n=1e7, d=64, k=256, 256 Gaussians); one step = one full initialisation; the unit
# is point-steps/s = n * (k - 1) / time (every step reads all rows once).
INIT = {"init_pp_c2": dict(n=10_000_000, d=64, k=256, method=2, center_range=20.0, sigma=(0.5, 2.0),
                           cpu_rows=200_000)}


# CSR rows (method lloyd_csr): n rows over d columns, each row storing the nnz
# columns of its cluster's support (256 clusters), values = center + noise;
# one step = one training call of `iters` Lloyd iterations from the first k rows.
CSR = {"csr_n1M_d2000_k256": dict(n=1_000_000, d=2000, k=256, nnz=20, iters=10, cpu_rows=1_000_000,
                                  cpu_iters=12)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_data(torch, n, d, k, center_range, sigma, anisotropic, seed, device):
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    centers = (torch.rand((k, d), generator=g, device=device, dtype=torch.float64) * 2 - 1) * center_range
    if anisotropic:
        scale = 10.0 ** (torch.rand((k, d), generator=g, device=device, dtype=torch.float64) * 2 - 1)
    else:
        s = torch.rand((k, 1), generator=g, device=device, dtype=torch.float64) * (sigma[1] - sigma[0]) + sigma[0]
        scale = s.expand(k, d)
    x = torch.empty((n, d), device=device, dtype=torch.float32)
    step = 1 << 20
    for r0 in range(0, n, step):
        r1 = min(n, r0 + step)
        lab = torch.randint(0, k, (r1 - r0,), generator=g, device=device)
        noise = torch.randn((r1 - r0, d), generator=g, device=device, dtype=torch.float32)
        x[r0:r1] = (centers[lab] + noise.double() * scale[lab]).float()
    perm = torch.randperm(n, generator=g, device=device)[:k]
    c0 = x[perm].clone()
    return x, c0


def cpu_model():
    """Model name of the host CPU (first 'model name' line of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return max(1, min(cores, 16))


def cpu_baseline(x_host, c0_host, iters, budget_s=15.0):
    """The reference's CPU Lloyd algorithm with a vectorised X.C^T kernel
    (oracle/kmeans_cpu_fast.cpp: DAAL's row blocks, per-thread partial sums merged
    in double, candidate lists for empty clusters), timed on a bounded sample of
    the workload (first rows, same initial centroids, one Lloyd iteration), on all
    host cores (at most 16) and on one."""
    from oracle import oracle as O

    cores = host_cores()
    n = x_host.shape[0]
    k = c0_host.shape[0]
    sample = min(n, max(k, 65536))
    t0 = time.perf_counter()
    O.cpu_fast_lloyd(x_host[:sample], c0_host, 1, nthreads=cores, final=False)
    dt = time.perf_counter() - t0
    # size the sample for about budget_s seconds of work (one iteration)
    sample = int(min(n, max(sample, sample * budget_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    O.cpu_fast_lloyd(x_host[:sample], c0_host, 1, nthreads=cores, final=False)
    dt = time.perf_counter() - t0
    value = sample / dt
    # the same kernel on one core, on a sample of about 4 s of work
    sample1 = int(min(n, max(k, value / cores * 4.0)))
    t0 = time.perf_counter()
    O.cpu_fast_lloyd(x_host[:sample1], c0_host, 1, nthreads=1, final=False)
    dt1 = time.perf_counter() - t0
    return {"value": value, "unit": "point-iterations/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(), "isa": O.cpu_fast_isa(),
            "sample": "first %d rows of the workload, 1 Lloyd iteration, the reference's CPU "
                      "Lloyd restated with a vectorised %s X.C^T kernel (oracle/kmeans_cpu_fast.cpp, "
                      "g++ -O3; float32, %d threads, %.1f s)" % (sample, O.cpu_fast_isa(), cores, dt),
            "single_core": {"value": sample1 / dt1, "cores": 1,
                            "sample": "first %d rows, 1 iteration, %.1f s" % (sample1, dt1)}}


def make_batch(torch, wl, count, seed, device):
    """Instances of the batched workload as lists of device tensors (x_i, c0_i)."""
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    gd = torch.Generator(device=device)
    gd.manual_seed(seed)
    xs, cs = [], []
    lo, hi = math.log(wl["n_lo"]), math.log(wl["n_hi"])
    for i in range(count):
        n = int(math.exp(lo + (hi - lo) * float(torch.rand(1, generator=g))))
        d = wl["dims"][int(torch.randint(0, len(wl["dims"]), (1,), generator=g))]
        k = int(torch.randint(wl["k_lo"], wl["k_hi"] + 1, (1,), generator=g))
        centers = (torch.rand((k, d), generator=gd, device=device) * 2 - 1) * 10.0
        lab = torch.randint(0, k, (n,), generator=gd, device=device)
        x = centers[lab] + torch.randn((n, d), generator=gd, device=device)
        perm = torch.randperm(n, generator=gd, device=device)[:k]
        xs.append(x.contiguous())
        cs.append(x[perm].contiguous())
    return xs, cs


def run_batched(torch, args, wl):
    from onedal_amd.engine import Engine, KM_OPT_TIMING

    count = args.instances or wl["count"]
    iters = wl["iters"]
    dev = torch.device("cuda:0")
    log("generating %s: %d instances" % (args.workload, count))
    xs, cs = make_batch(torch, wl, count, 0, dev)
    torch.cuda.synchronize()
    rows = sum(x.shape[0] for x in xs)
    flops_pit = sum(2.0 * x.shape[0] * c.shape[0] * x.shape[1] for x, c in zip(xs, cs))  # per iteration
    log("rows %d, X %.1f GB" % (rows, sum(x.numel() for x in xs) * 4 / 1e9))
    eng = Engine(0)
    eng.set_option(KM_OPT_TIMING, 1)
    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        table = eng.batch_autotune()
        tune = {"seconds": round(time.perf_counter() - t0, 2), "table": table}
        log("tile autotune: %.2f s" % tune["seconds"])
    outs = ([torch.empty_like(c) for c in cs], [torch.empty(x.shape[0], dtype=torch.int32, device=dev)
                                                for x in xs])
    for w in range(args.warmup):
        eng.train_batch_device(xs, cs, iters, graph=True, outs=outs)
        log("warmup %d done" % w)
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    loop_ms = []
    for s in range(args.steps):
        _, _, obj, its = eng.train_batch_device(xs, cs, iters, graph=True, outs=outs)
        torch.cuda.synchronize()
        loop_ms.append(eng.stats()["iterations_ms"])
        log("step %d: loop %.1f ms" % (s, loop_ms[-1]))
    elapsed = time.perf_counter() - t_all
    pit = rows * iters * args.steps / elapsed
    whole = flops_pit * iters * args.steps / elapsed / 1e12
    # dominant stage: the assignment (all tile kinds of one iteration), timed with
    # HIP events in one eager (not graph-replayed) run of the same job
    log("eager timed run for the assignment stage ...")
    eng.train_batch_device(xs, cs, iters, graph=False, outs=outs)
    torch.cuda.synchronize()
    st = eng.stats()
    asg_ms = st["assign_kernel_ms"] / max(1, st["assign_launches"])
    achieved = flops_pit / (asg_ms * 1e-3) / 1e12
    # the same stage against HBM: one read of every row (4 d bytes) and its label
    # write (4 bytes) per iteration; small-k buckets are bound by this, not by flops
    bytes_pit = sum(x.shape[0] * (4.0 * x.shape[1] + 4.0) for x in xs)
    hbm_gbs = bytes_pit / (asg_ms * 1e-3) / 1e9
    if args.copy_calibration:
        log("copy calibration: 3 x 1 GiB device-to-device copies ...")
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        from oracle import oracle as O

        cores = host_cores()
        t0 = time.perf_counter()
        done = 0
        used = 0
        for x, c in zip(xs, cs):
            xh, ch = x.cpu().numpy(), c.cpu().numpy()
            O.cpu_fast_lloyd(xh, ch, 1, nthreads=cores, final=False)
            done += xh.shape[0]
            used += 1
            if time.perf_counter() - t0 > 15.0:
                break
        dt = time.perf_counter() - t0
        cpu = {"value": done / dt, "unit": "point-iterations/s", "cores": cores, "kind": "port",
               "cpu_model": cpu_model(), "isa": O.cpu_fast_isa(),
               "sample": "first %d instances of the batch (%d rows), 1 Lloyd iteration each, the "
                         "reference's CPU Lloyd with a vectorised X.C^T kernel (float32, %d threads, "
                         "%.1f s incl. copies)" % (used, done, cores, dt)}
    out = {
        "metric": METRIC,
        "value": pit,
        "unit": "point-iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "instances": count, "rows": rows, "iterations": iters,
                   "path": "batched assignment (exact-form VALU tiles or certified FP32 matrix "
                           "cores per tuned (d, k) bucket), the whole job one HIP graph",
                   "tiles": "autotuned per (d, k) bucket" if tune else "default by k"},
        "autotune": tune,
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2),
                     "peak": PEAK["fp32_mfma_tflops"], "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK["fp32_mfma_tflops"], 4), "traffic": None,
                     "kernel": "assignment stage of one iteration (all tile kinds; 2kd flops per "
                               "row; HIP events, eager run)",
                     "avg_stage_ms": round(asg_ms, 3),
                     "whole_iteration_tflops": round(whole, 2),
                     "hbm": {"algorithmic_bytes_per_stage": bytes_pit, "achieved_gbs": round(hbm_gbs, 1),
                             "peak_gbs": PEAK["hbm_gbs"], "frac": round(hbm_gbs / PEAK["hbm_gbs"], 4)}},
        "cpu_baseline": cpu,
        "loop_ms_per_step": sum(loop_ms) / len(loop_ms),
    }
    print(json.dumps(out), flush=True)


def _pct(v, q):
    if not v:
        return None
    v = sorted(v)
    return v[min(len(v) - 1, int(math.ceil(q / 100.0 * len(v))) - 1)]


def run_init(torch, args, wl):
    from onedal_amd.engine import Engine

    n, d, k = wl["n"], wl["d"], wl["k"]
    dev = torch.device("cuda:0")
    log("generating %s: n=%d d=%d k=%d" % (args.workload, n, d, k))
    x, _ = make_data(torch, n, d, k, wl["center_range"], wl["sigma"], False, 0, dev)
    torch.cuda.synchronize()
    eng = Engine(0)
    for w in range(args.warmup):
        eng.init_centroids(x, k, wl["method"], 777)
        torch.cuda.synchronize()
        log("warmup %d done" % w)
    times = []
    for s in range(args.steps):
        t0 = time.perf_counter()
        c = eng.init_centroids(x, k, wl["method"], 777 + s)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        log("step %d: %.3f s" % (s, times[-1]))
    elapsed = sum(times)
    steps_per_s = n * (k - 1) * args.steps / elapsed
    # every k-means++ step reads X once (n*d*4 B) and reads and writes D2 (16 B per row)
    bytes_per_step = n * d * 4 + 16 * n
    gbs = bytes_per_step * (k - 1) * args.steps / elapsed / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        from oracle import oracle as O

        rows = wl["cpu_rows"]
        xs = x[:rows].cpu().numpy()
        t0 = time.perf_counter()
        O.init_centroids(xs, k, wl["method"], 777)
        dt = time.perf_counter() - t0
        cpu = {"value": rows * (k - 1) / dt, "unit": "point-steps/s", "cores": 1, "kind": "port",
               "sample": "first %d rows, full k-means++ (k=%d), oracle restatement, one thread, %.1f s"
                         % (rows, k, dt)}
    out = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 data, f64 distances",
        "data": "synthetic",
        "config": {"workload": args.workload, "n": n, "d": d, "k": k, "method": "plus_plus_dense"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK["hbm_gbs"], "unit": "GB/s",
                     "frac": round(gbs / PEAK["hbm_gbs"], 4), "traffic": None,
                     "kernel": "whole k-means++ step (D2 update + chunk sums + pick): n*d*4 + 16n B "
                               "per step over the whole call time"},
        "cpu_baseline": cpu,
        "centroid_checksum": float(c.double().sum().item()),
    }
    print(json.dumps(out), flush=True)


def run_serving(torch, args, wl, rates=None, min_jobs=None, window_s=None, cpu_baseline_on=True):
    """The request-loop sweep; returns the result line (a dict)."""
    import queue
    import random
    import threading

    from onedal_amd.engine import Server

    dev = torch.device("cuda:0")
    log("generating %s pools: %d small datasets, %d large" % (args.workload, wl["pool_small"],
                                                              wl["pool_large"]))
    xs, cs = make_batch(torch, wl, wl["pool_small"], 1, dev)
    nl, dl, kl = wl["large"]
    large = [make_data(torch, nl, dl, kl, 10.0, (1.0, 1.0), False, 100 + i, dev)
             for i in range(wl["pool_large"])]
    torch.cuda.synchronize()
    log("pool X %.1f GB" % ((sum(x.numel() for x in xs) + sum(x.numel() for x, _ in large)) * 4 / 1e9))
    iters = wl["iters"]
    srv = Server(budget_bytes=int(wl["budget"]), autotune=not args.no_autotune)

    def run(rate, count, seed):
        rng = random.Random(seed)
        reqs = []
        t = 0.0
        for _ in range(count):
            t += rng.expovariate(rate)
            big = rng.random() < wl["large_frac"]
            fit = rng.random() < wl["fit_frac"]
            if big:
                x, c = large[rng.randrange(len(large))]
            else:
                i = rng.randrange(len(xs))
                x, c = xs[i], cs[i]
            reqs.append((t, big, fit, x, c))
        ids = queue.Queue()
        done = []
        t_last = [0.0]

        def waiter():
            while True:
                item = ids.get()
                if item is None:
                    return
                jid, big, fit, n, k, d = item
                r = srv.wait(jid)
                t_last[0] = time.perf_counter()
                done.append((big, fit, n, k, d, r["latency_ms"], r["queue_ms"], r["lane"]))

        th = threading.Thread(target=waiter)
        th.start()
        st0 = srv.stats()
        t0 = time.perf_counter()
        for (ta, big, fit, x, c) in reqs:
            now = time.perf_counter() - t0
            if ta > now:
                time.sleep(ta - now)
            if fit:
                jid = srv.submit_fit(x, c, iters, bf16=True)
            else:
                jid = srv.submit_infer(x, c, bf16=True)
            ids.put((jid, big, fit, x.shape[0], c.shape[0], x.shape[1]))
        t_sub = time.perf_counter() - t0
        ids.put(None)
        th.join()
        st1 = srv.stats()
        el = t_last[0] - t0
        pit = sum(n * (iters if fit else 1) for (_, fit, n, _, _, _, _, _) in done)
        # distance flops (2kd per row per assignment pass, final pass included) by
        # lane: the batched lane computes in FP32, the large-job lane on BF16 matrix cores
        fl = [0.0, 0.0]
        pl = [0, 0]
        for (_, fit, n, k, d, _, _, lane) in done:
            fl[lane] += 2.0 * n * k * d * (iters + 1 if fit else 1)
            pl[lane] += n * (iters if fit else 1)
        busy = fl[0] / (PEAK["fp32_mfma_tflops"] * 1e12) + fl[1] / (PEAK["bf16_mfma_tflops"] * 1e12)
        flops = fl[0] + fl[1]
        lat = [r[5] for r in done]
        small_lat = [r[5] for r in done if not r[0]]
        big_lat = [r[5] for r in done if r[0]]
        steps = st1["steps"] - st0["steps"]
        return {"offered_rate": rate, "jobs": count, "arrival_span_s": round(reqs[-1][0], 3),
                "submit_span_s": round(t_sub, 3), "elapsed_s": round(el, 3),
                "jobs_per_s": round(count / el, 1), "point_iters_per_s": pit / el,
                "tflops": round(flops / el / 1e12, 2),
                "peak_weighted_frac": round(busy / el, 4),
                "batched_lane": {"point_iters_per_s": pl[0] / el, "tflops": round(fl[0] / el / 1e12, 2)},
                "large_lane": {"point_iters_per_s": pl[1] / el, "tflops": round(fl[1] / el / 1e12, 2)},
                "p50_ms": _pct(lat, 50), "p99_ms": _pct(lat, 99),
                "p50_small_ms": _pct(small_lat, 50), "p99_small_ms": _pct(small_lat, 99),
                "p50_large_ms": _pct(big_lat, 50), "p99_large_ms": _pct(big_lat, 99),
                "mean_queue_ms": sum(r[6] for r in done) / len(done),
                "large_jobs": len(big_lat), "batched_steps": steps,
                "mean_step_jobs": round((st1["step_jobs"] - st0["step_jobs"]) / max(1, steps), 1),
                "max_step_jobs": st1["max_step_jobs"],
                "peak_admitted_gb": round(st1["peak_bytes_admitted"] / 1e9, 2)}

    log("warmup ...")
    run(200.0, 60, 12345)
    sweep = []
    if rates is None:
        rates = [float(r) for r in (args.rates.split(",") if args.rates else wl["rates"])]
    mj = wl["min_jobs"] if min_jobs is None else min_jobs
    win = wl["window_s"] if window_s is None else window_s
    for r in rates:
        count = int(min(wl["max_jobs"], max(mj, r * win)))
        res = run(r, count, int(r))
        sweep.append(res)
        log("rate %6.0f/s: %d jobs in %.2f s, %.0f jobs/s, %.3g pit/s, p50 %.1f ms, p99 %.1f ms, "
            "mean step jobs %.1f" % (r, count, res["elapsed_s"], res["jobs_per_s"],
                                      res["point_iters_per_s"], res["p50_ms"], res["p99_ms"],
                                      res["mean_step_jobs"]))
    srv.close()
    best = max(sweep, key=lambda z: z["point_iters_per_s"])
    cpu = None
    if cpu_baseline_on and not args.no_cpu_baseline:
        log("cpu baseline ...")
        from oracle import oracle as O

        cores = host_cores()
        t0 = time.perf_counter()
        done_rows = used = 0
        for x, c in zip(xs, cs):
            O.cpu_fast_lloyd(x.cpu().numpy(), c.cpu().numpy(), 1, nthreads=cores, final=False)
            done_rows += x.shape[0]
            used += 1
            if time.perf_counter() - t0 > 10.0:
                break
        dt = time.perf_counter() - t0
        cpu = {"value": done_rows / dt, "unit": "point-iterations/s", "cores": cores, "kind": "port",
               "cpu_model": cpu_model(), "isa": O.cpu_fast_isa(),
               "sample": "first %d small datasets of the pool (%d rows), 1 Lloyd iteration each, "
                         "the reference's CPU Lloyd with a vectorised X.C^T kernel (float32, %d "
                         "threads, %.1f s incl. copies)" % (used, done_rows, cores, dt)}
    out = {
        "metric": METRIC,
        "value": best["point_iters_per_s"],
        "unit": "point-iterations/s",
        "n_gpus": 1,
        "steps": len(sweep),
        "warmup": 1,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (batched lane), bf16+f32 (large-job lane)",
        "data": "synthetic",
        "config": {"workload": args.workload, "fit_fraction": wl["fit_frac"],
                   "large_fraction": wl["large_frac"], "fit_iterations": iters,
                   "budget_bytes": wl["budget"], "large_job": {"n": nl, "d": dl, "k": kl}},
        "latency": {"p50_ms_at_best": best["p50_ms"], "p99_ms_at_best": best["p99_ms"],
                    "rate_at_best": best["offered_rate"],
                    # the highest offered rate of the sweep still served with p99 <= 100 ms
                    "max_rate_p99_le_100ms": max([z["offered_rate"] for z in sweep if z["p99_ms"] <= 100.0],
                                                 default=None)},
        # peak of the lane mix: the flops each lane ran, each priced at its precision's
        # dense peak (batched lane FP32, large-job lane BF16)
        "roofline": {"bound": "mfma", "achieved": best["tflops"],
                     "peak": round(best["tflops"] / max(best["peak_weighted_frac"], 1e-12), 1),
                     "unit": "TFLOP/s", "frac": best["peak_weighted_frac"], "traffic": None,
                     "kernel": "whole request loop: 2kd flops per row per assignment pass; peak "
                               "weighted by lane (FP32 157.3 batched, BF16 2500 large-job)"},
        "cpu_baseline": cpu,
        "sweep": sweep,
    }
    del xs, cs, large
    torch.cuda.empty_cache()
    return out


def run_csr(torch, args, wl):
    from onedal_amd.engine import Engine, KM_OPT_TIMING

    n, d, k, m, iters = wl["n"], wl["d"], wl["k"], wl["nnz"], wl["iters"]
    dev = torch.device("cuda:0")
    log("generating %s: n=%d d=%d k=%d, %d entries per row" % (args.workload, n, d, k, m))
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    supports = torch.sort(torch.rand(k, d, generator=g, device=dev).argsort(dim=1)[:, :m], dim=1).values
    centers = torch.rand(k, d, generator=g, device=dev) * 8.0 + 2.0
    truth = torch.randint(0, k, (n,), generator=g, device=dev)
    cols = supports[truth]                                     # n x m, sorted per row
    vals = (centers[truth.unsqueeze(1), cols] +
            0.3 * torch.randn(n, m, generator=g, device=dev)).float().reshape(-1).contiguous()
    cols = cols.reshape(-1).to(torch.int64).contiguous()
    offs = torch.arange(0, n * m + 1, m, device=dev, dtype=torch.int64)
    # initial centroids: the first k rows, expanded
    c0 = torch.zeros(k, d, device=dev, dtype=torch.float32)
    c0.scatter_(1, cols[: k * m].reshape(k, m), vals[: k * m].reshape(k, m))
    torch.cuda.synchronize()
    eng = Engine(0)
    eng.set_option(KM_OPT_TIMING, 1)
    for w in range(args.warmup):
        eng.train_csr(vals, cols, offs, d, c0, iters, 0.0)
        torch.cuda.synchronize()
        log("warmup %d done" % w)
    times, asg_ms, launches = [], 0.0, 0
    for s_ in range(args.steps):
        t0 = time.perf_counter()
        c, labels, obj, it = eng.train_csr(vals, cols, offs, d, c0, iters, 0.0)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        st = eng.stats()
        asg_ms += st["assign_kernel_ms"]
        launches += st["assign_launches"]
        log("step %d: %.3f s, objective %.9e, assign %.3f ms/launch" % (
            s_, times[-1], obj, st["assign_kernel_ms"] / max(1, st["assign_launches"])))
    elapsed = sum(times)
    value = n * iters * args.steps / elapsed
    avg_ms = asg_ms / max(1, launches)
    flops = 2.0 * n * m * k  # one multiply and one add per stored entry and centroid
    tflops = flops / (avg_ms * 1e-3) / 1e12
    # the assignment's dominant traffic: per stored entry, one centroid column of k
    # values gathered through the caches (the transposed centroids, k*d*4 B, stay in L2)
    gather_gbs = n * m * k * 4 / (avg_ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        from oracle import oracle as O

        rows = wl["cpu_rows"]
        hv = vals[: rows * m].cpu().numpy()
        hc = cols[: rows * m].cpu().numpy()
        ho = offs[: rows + 1].cpu().numpy()
        xs = O.Csr(hv, hc, ho, d, 0)
        ci = wl["cpu_iters"]
        t0 = time.perf_counter()
        O.csr_lloyd(xs, c0.cpu().numpy(), ci, 0.0, nthreads=16)
        dt = time.perf_counter() - t0
        cpu = {"value": rows * ci / dt, "unit": "point-iterations/s", "cores": 16, "kind": "port",
               "sample": "first %d rows, %d Lloyd iterations plus the final assignment, oracle CSR "
                         "restatement, 16 threads, %.1f s" % (rows, ci, dt)}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "point-iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "n": n, "d": d, "k": k, "entries_per_row": m,
                   "iterations": iters, "method": "lloyd_csr"},
        "roofline": {"bound": "valu", "achieved": round(tflops, 2), "peak": PEAK["fp32_vector_tflops"],
                     "unit": "TFLOP/s", "frac": round(tflops / PEAK["fp32_vector_tflops"], 4), "traffic": None,
                     "kernel": "csr_assign_kernel: 2*nnz*k flops per launch (separate multiply and "
                               "add, so the vector units' ceiling is half the FMA peak)",
                     "avg_launch_ms": round(avg_ms, 4), "cache_gather_gbs": round(gather_gbs, 1)},
        "cpu_baseline": cpu,
        "objective": obj,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="precision_c3",
                    choices=sorted(WORKLOADS) + sorted(BATCHED) + sorted(SERVING) + sorted(INIT) +
                    sorted(CSR))
    ap.add_argument("--rates", default=None,
                    help="serving workloads: comma-separated offered rates in jobs/s")
    ap.add_argument("--instances", type=int, default=None,
                    help="batched workloads: number of instances (default: the workload's)")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the FP32-path and FP64-path comparison runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="precision_c3: skip the FP32-workload summary and the short serving "
                         "sweep reported beside the headline line")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched and serving workloads: keep the default tile per k instead of "
                         "autotuning the batched assignment's tiles first")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the run, copy a 1 GiB device buffer three times (a plain copy "
                         "of known size, to scale hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the engine has no multi-GPU path "
              "(rows would shard with one all-reduce of the FP64 sums per iteration); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.workload in BATCHED:
        return run_batched(torch, args, BATCHED[args.workload])
    if args.workload in INIT:
        return run_init(torch, args, INIT[args.workload])
    if args.workload in SERVING:
        print(json.dumps(run_serving(torch, args, SERVING[args.workload])), flush=True)
        return
    if args.workload in CSR:
        return run_csr(torch, args, CSR[args.workload])
    out = run_dense(torch, args, args.workload, args.steps, args.warmup,
                    compare=not args.no_compare, cpu_on=not args.no_cpu_baseline)
    if args.workload == "precision_c3" and not args.no_extras:
        # the other workloads of the default run, reported beside the headline line:
        # the plain FP32 workload (throughput, roofline, labels against the reference's
        # CPU algorithm at the final centroids) and a short serving sweep
        log("extra: fp32_c2 ...")
        c2 = run_dense(torch, args, "fp32_c2", 2, 1, compare=False, cpu_on=False, label_check=True)
        out["fp32_c2"] = {k_: c2[k_] for k_ in ("value", "unit", "ms_per_step", "steps", "warmup",
                                                "config", "roofline", "labels_vs_cpu_reference")}
        log("extra: serving sweep ...")
        sv = run_serving(torch, args, SERVING["serving_c5"], rates=[50.0, 200.0, 400.0],
                         min_jobs=150, window_s=1.5, cpu_baseline_on=False)
        out["serving_sweep"] = {
            "config": sv["config"],
            "note": "short sweep (at least 150 jobs or 1.5 s of arrivals per rate); the full "
                    "sweep is bench.py --workload serving_c5",
            "rates": [{k_: z[k_] for k_ in ("offered_rate", "jobs", "jobs_per_s",
                                            "point_iters_per_s", "p50_ms", "p99_ms",
                                            "p50_small_ms", "p99_small_ms", "p50_large_ms",
                                            "p99_large_ms", "mean_step_jobs")}
                      for z in sv["sweep"]]}
        out["extra_keys"] = ["fp32_c2", "serving_sweep"]
    print(json.dumps(out), flush=True)


def band_mask_fp64(x, c, labels_fp64=None):
    """Rows whose FP64 top-2 squared-distance gap is below eps (|x|^2 + |c_1|^2),
    eps = 4 d 2^-24: inside it float rounding may order the two nearest centroids
    either way.  x, c: float32 CUDA tensors; returns a bool CUDA tensor."""
    import torch

    n, d = x.shape
    eps = 4.0 * d * 2.0 ** -24
    cd = c.double()
    cn = (cd * cd).sum(1)
    out = torch.empty(n, dtype=torch.bool, device=x.device)
    step = 1 << 18
    for r0 in range(0, n, step):
        xd = x[r0:r0 + step].double()
        xn = (xd * xd).sum(1)
        dist = (xn[:, None] - 2.0 * (xd @ cd.T) + cn[None, :]).clamp_min(0.0)
        top2, idx = torch.topk(dist, 2, dim=1, largest=False)
        gap = top2[:, 1] - top2[:, 0]
        out[r0:r0 + step] = gap < eps * (xn + cn[idx[:, 0]])
    return out


def run_dense(torch, args, name, steps, warmup, compare=True, cpu_on=True, label_check=False):
    """One dense workload (WORKLOADS[name]): timed training calls; returns the
    result line (a dict)."""
    from onedal_amd.engine import Engine, KM_OPT_TIMING

    wl = WORKLOADS[name]
    n, d, k, iters = wl["n"], wl["d"], wl["k"], wl["iters"]
    dev = torch.device("cuda:0")
    log("generating %s: n=%d d=%d k=%d" % (name, n, d, k))
    x, c0 = make_data(torch, n, d, k, wl["center_range"], wl["sigma"], wl["anisotropic"], 0, dev)
    torch.cuda.synchronize()
    eng = Engine(0)
    eng.set_option(KM_OPT_TIMING, 1)
    c_out = torch.empty((k, d), dtype=torch.float32, device=dev)
    l_out = torch.empty(n, dtype=torch.int32, device=dev)
    for w in range(warmup):
        eng.train_device(x, c0, iters, 0.0, bf16=wl["bf16"], out=(c_out, l_out))
        log("warmup %d done" % w)
    torch.cuda.synchronize()
    times = []
    stats = []
    t_all = time.perf_counter()
    for s_ in range(steps):
        t0 = time.perf_counter()
        _, _, obj, it = eng.train_device(x, c0, iters, 0.0, bf16=wl["bf16"], out=(c_out, l_out))
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        stats.append(eng.stats())
        log("step %d: %.3f s, iterations %d, objective %.9e, re-scored rows cand/collect/full %d/%d/%d" % (
            s_, times[-1], it, obj, stats[-1]["refined_candidate_rows"],
            stats[-1]["refined_collect_rows"], stats[-1]["refined_full_rows"]))
    elapsed = time.perf_counter() - t_all
    pit = n * iters * steps / elapsed

    # roofline of the dominant kernel (assignment): algorithmic 2*k*d flops per row
    launches = sum(s["assign_launches"] for s in stats)
    kern_ms = sum(s["assign_kernel_ms"] for s in stats)
    avg_ms = kern_ms / max(1, launches)
    flops_per_launch = 2.0 * k * d * n
    achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
    peak = PEAK["bf16_mfma_tflops"] if wl["bf16"] else PEAK["fp32_mfma_tflops"]
    mt = MEASURED_TRAFFIC.get(name)
    roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": mt["bytes"] if mt else None,
            "kernel": "assign (fused distance GEMM + argmin)", "avg_launch_ms": round(avg_ms, 4)}
    if mt:
        roof["traffic_detail"] = {"unit": "bytes per launch", "read": mt["read"],
                                  "write": mt["write"], "source": mt["source"],
                                  "hbm_gbs_at_avg_launch": round(mt["bytes"] / (avg_ms * 1e-3) / 1e9, 1)}
    if wl["bf16"]:
        # what a pure BF16 MFMA loop reaches on this device right now (the shader
        # clock under matrix-core load sits below the one the published peak assumes)
        ceil_tf, ceil_ghz = eng.measure_bf16_mfma_peak()
        roof["measured_mfma_ceiling"] = {"tflops": round(ceil_tf, 1), "clock_ghz": round(ceil_ghz, 3),
                                         "frac": round(achieved / ceil_tf, 4),
                                         "probe": "km_measure_bf16_mfma_peak: 4 independent "
                                                  "32x32x16 BF16 MFMA chains per wave, 8 waves "
                                                  "per CU, no memory traffic"}

    comparison = None
    if wl["bf16"] and compare:
        # same data and init on the plain FP32 path: throughput ratio and label identity
        log("fp32 path on the same workload ...")
        c32 = torch.empty_like(c_out)
        l32 = torch.empty_like(l_out)
        t0 = time.perf_counter()
        _, _, obj32, _ = eng.train_device(x, c0, iters, 0.0, bf16=False, out=(c32, l32))
        torch.cuda.synchronize()
        t32 = time.perf_counter() - t0
        st32 = eng.stats()
        same = bool(torch.equal(l32, l_out)) and bool(torch.equal(c32, c_out))
        comparison = {"fp32_path_value": n * iters / t32, "fp32_path_ms_per_step": 1e3 * t32,
                      "speedup_vs_fp32": (n * iters / t32 and pit / (n * iters / t32)),
                      "labels_and_centroids_identical_to_fp32": same,
                      "fp32_assign_avg_launch_ms": st32["assign_kernel_ms"] / max(1, st32["assign_launches"]),
                      "objective_bf16": obj, "objective_fp32": obj32}
        del c32, l32
        # FP64 Lloyd on the same rows and initial centroids (double data): the
        # engine's FP64 path computes the oracle's FP64 restatement bit for bit
        # (tests/test_gpu_parity.py), so it stands for the FP64 oracle at full size
        log("fp64 path on the same workload ...")
        xd = x.double()
        cd = c0.double()
        cdo = torch.empty_like(cd)
        ldo = torch.empty_like(l_out)
        t0 = time.perf_counter()
        _, _, obj64, _ = eng.train_device(xd, cd, iters, 0.0, out=(cdo, ldo))
        torch.cuda.synchronize()
        comparison["fp64_path_ms_per_step"] = 1e3 * (time.perf_counter() - t0)
        comparison["objective_fp64"] = obj64
        comparison["inertia_rel_err_vs_fp64"] = abs(obj - obj64) / obj64
        comparison["label_agreement_vs_fp64"] = float((ldo == l_out).float().mean().item())
        comparison["label_mismatches_vs_fp64"] = int((ldo != l_out).sum().item())
        # centroid error: per centroid, |c - c64| / |c64| (L2), and the largest
        # element difference relative to the largest centroid element
        dc = c_out.double() - cdo
        rel = dc.norm(dim=1) / cdo.norm(dim=1).clamp_min(1e-300)
        comparison["centroid_max_rel_err_vs_fp64"] = float(rel.max().item())
        comparison["centroid_max_abs_err_over_max_abs_vs_fp64"] = float(
            (dc.abs().max() / cdo.abs().max()).item())
        comparison["fp64_reference"] = ("the engine's FP64 path on float64 copies of the same rows "
                                        "and initial centroids (bit-identical to the oracle's FP64 "
                                        "restatement)")
        del xd, cd, cdo, ldo, dc
        torch.cuda.empty_cache()

    labels_check = None
    if label_check:
        # labels against the reference's CPU algorithm at the final centroids, on the
        # first rows: the oracle's DAAL restatement (float32, 16 threads) assigns the
        # same rows to the same centroids; differences must lie inside the float band
        from oracle import oracle as O

        m = min(n, 1_000_000)
        xs = x[:m].cpu().numpy()
        t0 = time.perf_counter()
        _, dl, _, _ = O.daal_lloyd(xs, c_out.cpu().numpy(), 0, 0.0, stop_rule=2, block_size=512,
                                   nthreads=host_cores())
        t_cpu = time.perf_counter() - t0
        dl = torch.from_numpy(dl).to(dev)
        band = band_mask_fp64(x[:m], c_out)
        mism = dl != l_out[:m]
        labels_check = {"rows": m, "mismatches": int(mism.sum().item()),
                        "mismatches_outside_band": int((mism & ~band).sum().item()),
                        "rows_inside_band": int(band.sum().item()),
                        "band": "FP64 top-2 gap < 4 d 2^-24 (|x|^2 + |c_1|^2)",
                        "reference": "oracle DAAL-Lloyd restatement (float32), assignment at the "
                                     "GPU's final centroids, %.1f s on %d threads" % (t_cpu, host_cores())}
        del xs, dl, band, mism

    cpu = None
    if cpu_on:
        log("cpu baseline ...")
        xs = x[: min(n, 4_000_000)].cpu().numpy()
        cpu = cpu_baseline(xs, c0.cpu().numpy(), iters, budget_s=15.0)
        del xs

    if args.copy_calibration:
        log("copy calibration: 3 x 1 GiB device-to-device copies ...")
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    out = {
        "metric": METRIC,
        "value": pit,
        "unit": "point-iterations/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": 1e3 * elapsed / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16+f32" if wl["bf16"] else "f32",
        "data": "synthetic",
        "config": {"workload": name, "n": n, "d": d, "k": k, "iterations": iters,
                   "path": "bf16+refine" if wl["bf16"] else "fp32",
                   "init": "random distinct rows (seed 0)"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "loop_ms_per_step": sum(s["iterations_ms"] for s in stats) / len(stats),
        "refined_rows_per_step": {k_: stats[-1][k_] for k_ in ("refined_candidate_rows",
                                                                 "refined_collect_rows",
                                                                 "refined_full_rows")},
        "comparison": comparison,
        "labels_vs_cpu_reference": labels_check,
    }
    eng.close()
    del x, c0, c_out, l_out
    torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    main()
