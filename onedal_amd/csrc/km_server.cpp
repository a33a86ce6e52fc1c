// km_server.cpp -- serving: a request loop over one device.
//
// The reference serves one train()/infer() call at a time
// (cpp/oneapi/dal/algo/kmeans/backend/gpu/train_kernel_lloyd_dense_dpc.cpp:76-207,
// infer_kernel_dpc.cpp:36-107).  Here jobs are submitted from any thread and run
// side by side:
//   admission  takes jobs in submission order, reserves each job's HBM footprint
//              against the budget (the head of the queue waits until it fits),
//              stages host-memory jobs into device blocks and hands the job to a
//              lane;
//   batched    continuous batching over km_batch_step_f32: every step iterates
//              all active fit jobs of the lane and, in the same step, finishes the
//              fit jobs whose iterations are done and every infer job; jobs
//              admitted meanwhile join at the next step;
//   large      jobs with n*k*d >= large_job_work, one at a time on a context (and
//              stream) of their own through km_lloyd_train_f32 / _infer_f32, so
//              their kernels overlap the batched lane's on the device.
// Both lanes compute the canonical arithmetic of the single-instance engine, so a
// job's labels, centroids and objective do not depend on what else is running.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/onedal_amd/kmeans_hip.h"
#include "km_args.hpp"

namespace kms {

using clock_type = std::chrono::steady_clock;

thread_local std::string t_last_error;

static km_status fail(km_status st, const std::string& msg) {
    t_last_error = msg;
    return st;
}

constexpr int64_t kDefaultLargeWork = int64_t(1) << 34;
constexpr int kDefaultMaxJobs = 4096;
constexpr int kMaxJobsLimit = 65536;        // sort keys: 16 bits of instance, 16 of label
constexpr int64_t kBatchedMaxK = 65536;     // larger k takes the large-job lane
constexpr int64_t kMaxChunk = 4;            // iterations per step while nothing waits
constexpr int64_t kRowsPerStep = (int64_t(1) << 31) - 1;

// Device blocks of staged host-memory jobs.  hipFree waits for the whole device,
// so freed blocks are kept and reused (the smallest cached block of at least the
// size and at most twice it); cached blocks are freed only to stay under the cap
// or when an allocation fails.
class block_cache {
public:
    void set_cap(int64_t cap) { cap_ = cap; }

    void* get(size_t bytes) {
        bytes = round_up(bytes);
        std::lock_guard<std::mutex> lk(mu_);
        auto it = free_.lower_bound(bytes);
        if (it != free_.end() && it->first <= 2 * bytes) {
            void* p = it->second;
            free_.erase(it);
            return p;
        }
        while (!free_.empty() && held_ + int64_t(bytes) > cap_) drop_largest();
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            while (!free_.empty()) drop_largest();
            if (hipMalloc(&p, bytes) != hipSuccess) {
                (void)hipGetLastError();
                return nullptr;
            }
        }
        held_ += int64_t(bytes);
        size_[p] = bytes;
        return p;
    }

    void put(void* p) {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = size_.find(p);
        if (it != size_.end()) free_.emplace(it->second, p);
    }

    // every block, cached or not (server teardown, after the lanes have stopped)
    void release_all() {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto& kv : size_) (void)hipFree(kv.first);
        size_.clear();
        free_.clear();
        held_ = 0;
    }

private:
    static size_t round_up(size_t b) {
        const size_t g = b >= (size_t(2) << 20) ? (size_t(2) << 20) : (size_t(64) << 10);
        return (std::max<size_t>(b, 1) + g - 1) / g * g;
    }
    void drop_largest() {
        auto last = std::prev(free_.end());
        (void)hipFree(last->second);
        held_ -= int64_t(last->first);
        size_.erase(last->second);
        free_.erase(last);
    }

    std::mutex mu_;
    std::multimap<size_t, void*> free_;
    std::unordered_map<void*, size_t> size_;
    int64_t held_ = 0;
    int64_t cap_ = 0;
};

struct job {
    int64_t id = 0;
    bool fit = false;
    bool host = false;  // caller buffers are host memory, staged through device blocks
    uint32_t flags = 0;
    int lane = 0;
    int64_t n = 0, d = 0, k = 0, max_iter = 0;
    int64_t footprint = 0;
    // caller buffers
    const float* X_user = nullptr;
    const float* C_user = nullptr;  // C0 (fit) or the model centroids (infer)
    float* Cout_user = nullptr;
    int32_t* labels_user = nullptr;
    // device views the kernels use
    const float* X = nullptr;
    const float* Cin = nullptr;
    float* Cstate = nullptr;  // fit: current centroids (the caller's C_out for device jobs)
    int32_t* labels = nullptr;
    std::vector<void*> blocks;
    int64_t iters_done = 0;
    // result
    bool done = false;
    int waiters = 0;
    km_status status = KM_OK;
    std::string error;
    double objective = 0.0;
    int64_t iterations = 0;
    clock_type::time_point t_submit, t_admit, t_done;
};

// HBM a job holds while admitted: its data, labels and centroids plus an estimate
// of the lane's workspace per row and per centroid element.
static int64_t footprint(bool fit, int lane, int64_t n, int64_t d, int64_t k) {
    const int64_t x = 4 * n * d, lab = 4 * n, cen = 4 * k * d;
    int64_t ws;
    if (lane == 0) {
        // concatenated labels, two copies of the sort keys, row ids, d64, sort
        // scratch; FP64 chunk partial sums; two concatenated centroid copies
        ws = fit ? 40 * n + 8 * d * (n / 256 + k + 1) + 2 * cen : cen + n;
    } else {
        // padded BF16 / FP32 row copies, row norms, lists, sort buffers
        int64_t dpad = 32;
        while (dpad < d) dpad <<= 1;
        ws = 4 * dpad * n + 64 * n + 8 * d * (n / 256 + k + 1) + 8 * cen;
    }
    return x + lab + 2 * cen + ws;
}

}  // namespace kms

struct km_server {
    km_server_config cfg{};
    int64_t large_work = 0;
    km_ctx* ctx_batch = nullptr;
    km_ctx* ctx_large = nullptr;
    hipStream_t stage = nullptr;      // admission uploads
    hipStream_t out_batch = nullptr;  // result downloads of each lane
    hipStream_t out_large = nullptr;
    kms::block_cache cache;

    std::mutex mu;
    std::condition_variable cv_admit, cv_batch, cv_large, cv_done;
    std::deque<kms::job*> pending, ready_batch, ready_large;
    std::unordered_map<int64_t, std::unique_ptr<kms::job>> jobs;
    int64_t next_id = 1;
    int64_t bytes_admitted = 0;
    bool stopping = false;
    bool admission_done = false;
    km_server_stats st{};
    std::thread t_admit, t_batch, t_large;
};

namespace kms {

static double ms_between(clock_type::time_point a, clock_type::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// A finished job: blocks back to the cache, budget released, waiters woken.
static void complete(km_server* s, job* j, km_status st, const std::string& msg, double obj,
                     int64_t iters) {
    for (void* b : j->blocks) s->cache.put(b);
    j->blocks.clear();
    std::lock_guard<std::mutex> lk(s->mu);
    j->status = st;
    j->error = msg;
    j->objective = obj;
    j->iterations = iters;
    j->t_done = clock_type::now();
    j->done = true;
    s->bytes_admitted -= j->footprint;
    if (st == KM_OK)
        s->st.completed++;
    else
        s->st.failed++;
    s->cv_done.notify_all();
    s->cv_admit.notify_one();
}

static bool hip_ok(hipError_t e, std::string& msg, const char* what) {
    if (e == hipSuccess) return true;
    (void)hipGetLastError();
    msg = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

// Device views of an admitted job; host jobs get device blocks and their inputs
// uploaded.
static km_status stage_job(km_server* s, job* j, std::string& msg) {
    if (!j->host) {
        j->X = j->X_user;
        j->Cin = j->C_user;
        j->Cstate = j->fit ? j->Cout_user : nullptr;
        j->labels = j->labels_user;
        return KM_OK;
    }
    const size_t xb = size_t(4) * j->n * j->d, cb = size_t(4) * j->k * j->d, lb = size_t(4) * j->n;
    const size_t sizes[4] = {xb, cb, j->fit ? cb : 0, lb};
    void* p[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int i = 0; i < 4; ++i) {
        if (sizes[i] == 0) continue;
        p[i] = s->cache.get(sizes[i]);
        if (!p[i]) {
            msg = "device allocation of " + std::to_string(sizes[i]) + " bytes failed";
            return KM_DEVICE_BAD_ALLOC;
        }
        j->blocks.push_back(p[i]);
    }
    j->X = static_cast<const float*>(p[0]);
    j->Cin = static_cast<const float*>(p[1]);
    j->Cstate = static_cast<float*>(p[2]);
    j->labels = static_cast<int32_t*>(p[3]);
    if (!hip_ok(hipMemcpyAsync(p[0], j->X_user, xb, hipMemcpyHostToDevice, s->stage), msg, "upload") ||
        !hip_ok(hipMemcpyAsync(p[1], j->C_user, cb, hipMemcpyHostToDevice, s->stage), msg, "upload") ||
        !hip_ok(hipStreamSynchronize(s->stage), msg, "upload"))
        return KM_INTERNAL_ERROR;
    return KM_OK;
}

// Results of a host job back into the caller's buffers.
static km_status download(hipStream_t st, job* j, std::string& msg) {
    if (!j->host) return KM_OK;
    if (!hip_ok(hipMemcpyAsync(j->labels_user, j->labels, size_t(4) * j->n, hipMemcpyDeviceToHost, st),
                msg, "download"))
        return KM_INTERNAL_ERROR;
    if (j->fit && !hip_ok(hipMemcpyAsync(j->Cout_user, j->Cstate, size_t(4) * j->k * j->d,
                                         hipMemcpyDeviceToHost, st),
                          msg, "download"))
        return KM_INTERNAL_ERROR;
    if (!hip_ok(hipStreamSynchronize(st), msg, "download")) return KM_INTERNAL_ERROR;
    return KM_OK;
}

static void admission_loop(km_server* s) {
    (void)hipSetDevice(s->cfg.device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_admit.wait(lk, [&] {
            return (!s->pending.empty() &&
                    s->bytes_admitted + s->pending.front()->footprint <= s->cfg.hbm_budget_bytes) ||
                   (s->stopping && s->pending.empty());
        });
        if (s->pending.empty()) break;
        job* j = s->pending.front();
        s->pending.pop_front();
        s->bytes_admitted += j->footprint;
        s->st.peak_bytes_admitted = std::max(s->st.peak_bytes_admitted, s->bytes_admitted);
        j->t_admit = clock_type::now();
        lk.unlock();
        std::string msg;
        const km_status st = stage_job(s, j, msg);
        if (st != KM_OK) {
            complete(s, j, st, msg, 0.0, 0);
            lk.lock();
            continue;
        }
        lk.lock();
        if (j->lane == 0) {
            s->ready_batch.push_back(j);
            s->cv_batch.notify_one();
        } else {
            s->ready_large.push_back(j);
            s->cv_large.notify_one();
        }
    }
    s->admission_done = true;
    s->cv_batch.notify_all();
    s->cv_large.notify_all();
}

static void batch_loop(km_server* s) {
    (void)hipSetDevice(s->cfg.device);
    const int max_jobs = s->cfg.max_batch_jobs;
    std::vector<job*> active, it_jobs, fin_jobs;
    std::vector<km_batch_instance_f32> it_set, fin_set;
    std::vector<double> obj;
    for (;;) {
        bool waiting = false;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_batch.wait(lk, [&] {
                return !active.empty() || !s->ready_batch.empty() || s->admission_done;
            });
            int64_t rows = 0;
            for (job* j : active) rows += j->n;
            while (!s->ready_batch.empty() && int(active.size()) < max_jobs) {
                job* j = s->ready_batch.front();
                if (rows + j->n > kRowsPerStep) break;
                rows += j->n;
                active.push_back(j);
                s->ready_batch.pop_front();
            }
            if (active.empty()) {
                if (s->admission_done && s->ready_batch.empty()) break;
                continue;
            }
            waiting = !s->ready_batch.empty() || !s->pending.empty();
        }
        // step composition: iterate every active fit job with iterations left, by
        // as many iterations as the closest one still needs (one while jobs wait)
        int64_t chunk = waiting ? 1 : kMaxChunk;
        for (job* j : active)
            if (j->fit && j->iters_done < j->max_iter) chunk = std::min(chunk, j->max_iter - j->iters_done);
        it_set.clear();
        fin_set.clear();
        it_jobs.clear();
        fin_jobs.clear();
        for (job* j : active) {
            const bool iter = j->fit && j->iters_done < j->max_iter;
            if (iter) {
                it_set.push_back(km_batch_instance_f32{j->X, j->iters_done > 0 ? j->Cstate : j->Cin,
                                                       j->Cstate, j->labels, j->n, j->d, j->k});
                it_jobs.push_back(j);
            }
            const int64_t after = j->iters_done + (iter ? chunk : 0);
            if (!j->fit || after == j->max_iter) {
                const float* c = (j->fit && after > 0) ? j->Cstate : j->Cin;
                fin_set.push_back(km_batch_instance_f32{j->X, c, j->fit ? j->Cstate : nullptr, j->labels,
                                                        j->n, j->d, j->k});
                fin_jobs.push_back(j);
            }
        }
        obj.assign(fin_set.size(), 0.0);
        const km_status st =
            km_batch_step_f32(s->ctx_batch, it_set.data(), int64_t(it_set.size()), chunk, fin_set.data(),
                              int64_t(fin_set.size()), 0, obj.data());
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->st.steps++;
            s->st.step_jobs += int64_t(active.size());
            s->st.max_step_jobs = std::max<int64_t>(s->st.max_step_jobs, int64_t(active.size()));
        }
        if (st != KM_OK) {
            // every active job took part in the step: all of them fail
            const std::string msg = km_last_error(s->ctx_batch);
            for (job* j : active) complete(s, j, st, msg, 0.0, 0);
            active.clear();
            continue;
        }
        for (job* j : it_jobs) j->iters_done += chunk;
        for (size_t i = 0; i < fin_jobs.size(); ++i) {
            job* j = fin_jobs[i];
            std::string msg;
            const km_status ds = download(s->out_batch, j, msg);
            complete(s, j, ds, msg, obj[i], j->fit ? j->max_iter : 0);
        }
        if (!fin_jobs.empty()) {
            std::vector<job*> keep;
            keep.reserve(active.size());
            for (job* j : active)
                if (std::find(fin_jobs.begin(), fin_jobs.end(), j) == fin_jobs.end()) keep.push_back(j);
            active.swap(keep);
        }
    }
}

static void large_loop(km_server* s) {
    (void)hipSetDevice(s->cfg.device);
    for (;;) {
        job* j = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_large.wait(lk, [&] { return !s->ready_large.empty() || s->admission_done; });
            if (s->ready_large.empty()) break;
            j = s->ready_large.front();
            s->ready_large.pop_front();
            s->st.large_jobs++;
        }
        const uint32_t fl = KM_FLAG_DEVICE_POINTERS | (j->flags & KM_FLAG_BF16_REFINE);
        double obj = 0.0;
        int64_t iters = 0;
        km_status st;
        if (j->fit)
            st = km_lloyd_train_f32(s->ctx_large, j->X, j->n, j->d, j->Cin, j->k, j->max_iter, 0.0, fl,
                                    j->Cstate, j->labels, &obj, &iters);
        else
            st = km_lloyd_infer_f32(s->ctx_large, j->X, j->n, j->d, j->Cin, j->k,
                                    fl | KM_FLAG_EXACT_OBJECTIVE, j->labels, &obj);
        std::string msg;
        if (st != KM_OK)
            msg = km_last_error(s->ctx_large);
        else
            st = download(s->out_large, j, msg);
        complete(s, j, st, msg, obj, iters);
    }
}

static km_status submit(km_server* s, bool fit, const float* X, int64_t n, int64_t d, const float* C,
                        int64_t k, int64_t max_iter, uint32_t flags, float* C_out, int32_t* labels,
                        int64_t* job_out) {
    if (!s || !job_out) return fail(KM_INVALID_ARGUMENT, "server and job id output are required");
    try {
        if (fit)
            km::check_train_args(X, n, d, C, k, max_iter, 0.0, C_out, labels);
        else
            km::check_infer_args(X, n, d, C, k, 0, labels, nullptr);
        if (flags & ~uint32_t(KM_FLAG_DEVICE_POINTERS | KM_FLAG_BF16_REFINE))
            return fail(KM_INVALID_ARGUMENT, "unsupported flags for a server job");
        const bool large = (s->large_work >= 0 &&
                            double(n) * double(k) * double(d) >= double(s->large_work)) ||
                           k > kBatchedMaxK;
        auto j = std::make_unique<job>();
        j->fit = fit;
        j->host = (flags & KM_FLAG_DEVICE_POINTERS) == 0;
        j->flags = flags;
        j->lane = large ? 1 : 0;
        j->n = n;
        j->d = d;
        j->k = k;
        j->max_iter = fit ? max_iter : 0;
        j->footprint = footprint(fit, j->lane, n, d, k);
        j->X_user = X;
        j->C_user = C;
        j->Cout_user = fit ? C_out : nullptr;
        j->labels_user = labels;
        if (j->footprint > s->cfg.hbm_budget_bytes)
            return fail(KM_DEVICE_BAD_ALLOC, "job footprint of " + std::to_string(j->footprint) +
                                                 " bytes exceeds the HBM budget of " +
                                                 std::to_string(s->cfg.hbm_budget_bytes) + " bytes");
        j->t_submit = clock_type::now();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->stopping) return fail(KM_INVALID_ARGUMENT, "the server is shutting down");
            j->id = s->next_id++;
            *job_out = j->id;
            s->pending.push_back(j.get());
            s->jobs.emplace(j->id, std::move(j));
            s->st.submitted++;
        }
        s->cv_admit.notify_one();
        return KM_OK;
    } catch (const km::status_error& e) {
        return fail(e.status, e.what());
    } catch (const std::bad_alloc&) {
        return fail(KM_HOST_BAD_ALLOC, "host allocation failed");
    }
}

static void teardown(km_server* s) {
    if (s->ctx_batch) (void)km_context_destroy(s->ctx_batch);
    if (s->ctx_large) (void)km_context_destroy(s->ctx_large);
    (void)hipSetDevice(s->cfg.device);
    for (hipStream_t st : {s->stage, s->out_batch, s->out_large})
        if (st) (void)hipStreamDestroy(st);
    s->cache.release_all();
    delete s;
}

}  // namespace kms

extern "C" {

const char* km_server_last_error(void) { return kms::t_last_error.c_str(); }

km_status km_server_create(const km_server_config* cfg, km_server** out) {
    using namespace kms;
    if (!cfg || !out) return fail(KM_INVALID_ARGUMENT, "configuration and output are required");
    *out = nullptr;
    if (cfg->hbm_budget_bytes <= 0) return fail(KM_DOMAIN_ERROR, "HBM budget must be positive");
    if (cfg->max_batch_jobs < 0 || cfg->max_batch_jobs > kMaxJobsLimit)
        return fail(KM_DOMAIN_ERROR, "max_batch_jobs must be in [0, 65536]");
    km_server* s = new (std::nothrow) km_server();
    if (!s) return fail(KM_HOST_BAD_ALLOC, "host allocation failed");
    s->cfg = *cfg;
    if (s->cfg.max_batch_jobs == 0) s->cfg.max_batch_jobs = kDefaultMaxJobs;
    s->large_work = cfg->large_job_work == 0 ? kDefaultLargeWork : cfg->large_job_work;
    s->cache.set_cap(cfg->hbm_budget_bytes);
    km_status st = km_context_create(cfg->device, &s->ctx_batch);
    if (st == KM_OK) st = km_context_create(cfg->device, &s->ctx_large);
    if (st != KM_OK) {
        teardown(s);
        return fail(st, "no usable HIP device " + std::to_string(cfg->device));
    }
    if (cfg->autotune && (st = km_batch_autotune(s->ctx_batch, nullptr, 0)) != KM_OK) {
        const std::string msg = km_last_error(s->ctx_batch);
        teardown(s);
        return fail(st, "tile autotuning failed: " + msg);
    }
    if (hipSetDevice(cfg->device) != hipSuccess ||
        hipStreamCreateWithFlags(&s->stage, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s->out_batch, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s->out_large, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        teardown(s);
        return fail(KM_INTERNAL_ERROR, "stream creation failed");
    }
    try {
        s->t_admit = std::thread(admission_loop, s);
        s->t_batch = std::thread(batch_loop, s);
        s->t_large = std::thread(large_loop, s);
    } catch (const std::exception& e) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stopping = true;
        }
        s->cv_admit.notify_all();
        if (s->t_admit.joinable()) s->t_admit.join();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->admission_done = true;
        }
        s->cv_batch.notify_all();
        s->cv_large.notify_all();
        if (s->t_batch.joinable()) s->t_batch.join();
        if (s->t_large.joinable()) s->t_large.join();
        teardown(s);
        return fail(KM_INTERNAL_ERROR, std::string("thread creation failed: ") + e.what());
    }
    *out = s;
    return KM_OK;
}

km_status km_server_destroy(km_server* s) {
    if (!s) return kms::fail(KM_INVALID_ARGUMENT, "null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_admit.notify_all();
    s->t_admit.join();
    s->t_batch.join();
    s->t_large.join();
    kms::teardown(s);
    return KM_OK;
}

km_status km_server_submit_fit_f32(km_server* s, const float* X, int64_t n, int64_t d,
                                   const float* C0, int64_t k, int64_t max_iter, uint32_t flags,
                                   float* C_out, int32_t* labels_out, int64_t* job_out) {
    return kms::submit(s, true, X, n, d, C0, k, max_iter, flags, C_out, labels_out, job_out);
}

km_status km_server_submit_infer_f32(km_server* s, const float* X, int64_t n, int64_t d,
                                     const float* C, int64_t k, uint32_t flags, int32_t* labels_out,
                                     int64_t* job_out) {
    return kms::submit(s, false, X, n, d, C, k, 0, flags, nullptr, labels_out, job_out);
}

km_status km_server_wait(km_server* s, int64_t id, int64_t timeout_ms, km_job_result* out) {
    using namespace kms;
    if (!s || !out) return fail(KM_INVALID_ARGUMENT, "server and result output are required");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return fail(KM_INVALID_ARGUMENT, "unknown job id " + std::to_string(id));
    job* j = it->second.get();
    j->waiters++;
    auto ready = [&] { return j->done; };
    bool done = true;
    if (timeout_ms < 0)
        s->cv_done.wait(lk, ready);
    else
        done = s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    j->waiters--;
    if (!done) return KM_PENDING;
    out->iterations = j->iterations;
    out->objective = j->objective;
    out->queue_ms = ms_between(j->t_submit, j->t_admit);
    out->latency_ms = ms_between(j->t_submit, j->t_done);
    out->lane = j->lane;
    out->reserved = 0;
    const km_status st = j->status;
    if (st != KM_OK) t_last_error = j->error;
    if (j->waiters == 0) s->jobs.erase(id);
    return st;
}

km_status km_server_get_stats(km_server* s, km_server_stats* out) {
    if (!s || !out) return kms::fail(KM_INVALID_ARGUMENT, "server and stats output are required");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->queued = int64_t(s->pending.size());
    out->bytes_admitted = s->bytes_admitted;
    return KM_OK;
}

}  // extern "C"
