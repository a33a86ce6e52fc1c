/* kmeans_hip.h -- C ABI of the MI355X k-means Lloyd engine (libkmeans_hip.so).
 *
 * Drop-in boundary for oneDAL's dense Lloyd GPU backend.  Each entry point
 * replaces one reference interface:
 *
 *   km_lloyd_train_f32 / _f64  <->  oneapi::dal::kmeans::backend::
 *       train_kernel_gpu<Float, method::lloyd_dense, task::clustering>::operator()
 *       (cpp/oneapi/dal/algo/kmeans/backend/gpu/train_kernel.hpp:24-29,
 *        implemented at train_kernel_lloyd_dense_dpc.cpp:76-207)
 *   km_lloyd_infer_f32 / _f64  <->  infer_kernel_gpu<Float, method::lloyd_dense, ...>
 *       (backend/gpu/infer_kernel.hpp:24-29, infer_kernel_dpc.cpp:36-107)
 *   km_lloyd_train_csr_f32 / _f64, km_lloyd_infer_csr_f32 / _f64  <->  the same
 *       kernels for method::lloyd_csr (train_kernel_lloyd_csr_dpc.cpp:76-242)
 *   km_context_create/destroy  <->  dal::backend::context_gpu + sycl::queue
 *       (cpp/oneapi/dal/backend/dispatcher.hpp:118-132): one context = one
 *       device + one in-order HIP stream.
 *
 * Plain pointers and sizes only; no exception crosses the ABI.  Status codes
 * map one-to-one onto the oneapi::dal exception types
 * (cpp/oneapi/dal/exceptions.hpp:28-100).  Preconditions are those of
 * cpp/oneapi/dal/algo/kmeans/detail/train_ops.hpp:40-61 and infer_ops.hpp:40-56
 * and are checked on the host, in the same order, before any launch.
 *
 * Arithmetic contract: labels are argmin_j of the exact-form distance
 * sum_f (x_f - c_jf)^2 in the input precision (ties to the lowest j); centroid
 * sums and the objective are accumulated in double in a fixed order, so results
 * do not depend on timing, and the FP32 and BF16 paths return identical labels.
 */
#ifndef ONEDAL_AMD_KMEANS_HIP_H
#define ONEDAL_AMD_KMEANS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct km_ctx km_ctx;

typedef enum {
    KM_OK = 0,
    KM_DOMAIN_ERROR = 1,      /* oneapi::dal::domain_error */
    KM_INVALID_ARGUMENT = 2,  /* oneapi::dal::invalid_argument */
    KM_DEVICE_BAD_ALLOC = 3,  /* oneapi::dal::device_bad_alloc */
    KM_HOST_BAD_ALLOC = 4,    /* oneapi::dal::host_bad_alloc */
    KM_UNIMPLEMENTED = 5,     /* oneapi::dal::unimplemented */
    KM_INTERNAL_ERROR = 6,    /* oneapi::dal::internal_error (HIP runtime failure) */
    KM_PENDING = 7            /* km_server_wait: the job has not finished yet (not an error) */
} km_status;

/* call flags */
#define KM_FLAG_EXACT_OBJECTIVE 0x1u /* infer: also return the objective
                                        (result_options::compute_exact_objective_function) */
#define KM_FLAG_BF16_REFINE 0x2u     /* float data, d <= 256: distances on BF16 matrix
                                        cores, near-ties re-scored in float; labels are
                                        identical to the FP32 path.  Ignored for double
                                        data and d > 256. */
#define KM_FLAG_DEVICE_POINTERS 0x4u /* X, C0/C and every output are device pointers
                                        on the context's device */
#define KM_FLAG_GRAPH 0x8u           /* batched training: capture one iteration as a
                                        HIP graph and replay it */

/* context options (km_context_set_option) */
#define KM_OPT_ASSIGN_PATH 1 /* 0 auto, 1 generic tile kernel, 2 FP32 matrix core */
#define KM_OPT_TIMING 2      /* 1: time the assignment kernels with HIP events */
#define KM_OPT_BATCH_TILE 3  /* batched assignment tile: 0 the context's table (tuned by
                                km_batch_autotune, else chosen by k), v + 1 force variant v
                                (0 <= v < KM_BATCH_TILE_VARIANTS) */

/* Tile variants of the batched assignment.  Variants 0..6 compute the exact
 * form on the vector units with CT centroids on the lanes of a row group and RW
 * rows on each lane, (CT, RW) = (64,8) (64,4) (32,8) (16,8) (8,8) (32,4) (16,4).
 * Variant 7 computes FP32 matrix-core distances with a proven error band and
 * re-scores in the exact form every row the band cannot settle; it applies to
 * instances with d <= 128 and k <= 128 (others take the default tile).  All
 * variants compute the same labels. */
#define KM_BATCH_TILE_VARIANTS 8
/* Buckets of the tile table: 6 * d_bucket + k_bucket, d in {<=8, <=16, <=32,
 * <=64, <=128, <=256, more}, k in {<=8, <=16, <=32, <=64, <=128, more}. */
#define KM_BATCH_TILE_BUCKETS 42

typedef struct {
    int64_t assign_launches;      /* main assignment kernel launches in the last call
                                     (batched training without KM_FLAG_GRAPH: assignment
                                     stages, one per iteration, KM_OPT_TIMING only) */
    double assign_kernel_ms;      /* their summed duration (KM_OPT_TIMING only) */
    double iterations_ms;         /* Lloyd loop wall time on the stream (KM_OPT_TIMING only) */
    int64_t refined_candidate_rows; /* rows settled by the candidate re-score */
    int64_t refined_full_rows;      /* rows re-scored against every centroid */
    int64_t empty_cluster_events;   /* iterations that refilled empty clusters */
    int32_t assign_path;            /* path used: 1 generic, 2 FP32 MFMA, 3 BF16 MFMA */
    int32_t reserved;
    int64_t refined_collect_rows;   /* BF16 path: rows whose candidates were re-collected */
} km_stats;

/* Create a context on HIP device `device` with its own stream. */
km_status km_context_create(int device, km_ctx** out);
km_status km_context_destroy(km_ctx* ctx);
/* Message of the last failed call on this context ("" if none). */
const char* km_last_error(const km_ctx* ctx);
km_status km_context_set_option(km_ctx* ctx, int option, int64_t value);
km_status km_get_stats(const km_ctx* ctx, km_stats* out);

/* Measured BF16 matrix-core ceiling of the context's device: a loop of
 * independent v_mfma_f32_32x32x16_bf16 chains (two workgroups of four waves per
 * compute unit, no memory traffic) runs `iters` iterations per wave;
 * *tflops_out = dense flops / elapsed time, *clock_ghz_out (may be NULL) = the
 * shader clock one wave saw over its loop.  A measurement aid for the
 * assignment's roofline, with no counterpart in the reference. */
km_status km_measure_bf16_mfma_peak(km_ctx* ctx, int64_t iters, double* tflops_out,
                                    double* clock_ghz_out);

/* Device memory on the context's device, for callers that keep data there (the
 * C++ surface's device tables).  km_memcpy copies synchronously on the context's
 * stream.  No counterpart in the reference, whose device tables hold SYCL USM
 * (sycl::malloc_device, queue.memcpy; table/homogen.hpp:120-160). */
#define KM_COPY_HOST_TO_DEVICE 1
#define KM_COPY_DEVICE_TO_HOST 2
#define KM_COPY_DEVICE_TO_DEVICE 3
km_status km_device_alloc(km_ctx* ctx, int64_t bytes, void** out);
km_status km_device_free(km_ctx* ctx, void* ptr);
km_status km_memcpy(km_ctx* ctx, void* dst, const void* src, int64_t bytes, int32_t kind);

/* Lloyd training.  X: n x d row-major; C0: k x d row-major initial centroids.
 * Outputs: C_out k x d, labels_out n (int32), *obj_out objective (sum of squared
 * distances to the final centroids), *iters_out iterations run.
 * Stop rule of the oneDAL GPU driver (train_kernel_lloyd_dense_dpc.cpp:147-168):
 * after iteration t, stop if acc_thr > 0 and obj_t + acc_thr > obj_{t-1}. */
km_status km_lloyd_train_f32(km_ctx* ctx, const float* X, int64_t n, int64_t d, const float* C0,
                             int64_t k, int64_t max_iter, double acc_thr, uint32_t flags,
                             float* C_out, int32_t* labels_out, double* obj_out,
                             int64_t* iters_out);
km_status km_lloyd_train_f64(km_ctx* ctx, const double* X, int64_t n, int64_t d, const double* C0,
                             int64_t k, int64_t max_iter, double acc_thr, uint32_t flags,
                             double* C_out, int32_t* labels_out, double* obj_out,
                             int64_t* iters_out);

/* Lloyd inference: labels of X against centroids C (k x d); objective when
 * KM_FLAG_EXACT_OBJECTIVE is set (obj_out may be NULL otherwise). */
km_status km_lloyd_infer_f32(km_ctx* ctx, const float* X, int64_t n, int64_t d, const float* C,
                             int64_t k, uint32_t flags, int32_t* labels_out, double* obj_out);
km_status km_lloyd_infer_f64(km_ctx* ctx, const double* X, int64_t n, int64_t d, const double* C,
                             int64_t k, uint32_t flags, int32_t* labels_out, double* obj_out);

/* km_lloyd_infer_f32 that allocates no memory.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory during a call (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * X, C and labels_out must be device pointers (KM_FLAG_DEVICE_POINTERS is
 * required); the results equal km_lloyd_infer_f32's bit for bit.  Workspace is a
 * fixed 64 MiB array per device, kept until the process ends; rows run through
 * it in blocks, and a k x d whose centroid-side workspace leaves no room for a
 * block of 4096 rows is refused with KM_INVALID_ARGUMENT.  All calls on one
 * device share that array: calls that overlap on one device (several contexts or
 * threads) race and must be serialised by the caller. */
km_status km_lloyd_infer_f32_noalloc(km_ctx* ctx, const float* X, int64_t n, int64_t d,
                                     const float* C, int64_t k, uint32_t flags, int32_t* labels_out,
                                     double* obj_out);

/* CSR data (method lloyd_csr).  values[nnz], col_indices[nnz], row_offsets[n + 1]
 * as csr_table holds them (table/csr.hpp), zero- or one-based (sparse_indexing);
 * column indices need not be sorted.  Offsets must run from the base to
 * base + nnz without decreasing and every column index must lie in
 * [base, base + d): otherwise KM_INVALID_ARGUMENT.  Distances are the expanded
 * form of the reference (assign_clusters, backend/gpu/kernels_csr_impl.hpp:118-187):
 * s_j = |c_j|^2 - 2 x.c_j over the stored entries in stored order, label = first
 * j with the smallest s_j, closest = s_label + |x|^2; the objective is the sum of
 * the closest distances (calc_objective_function, :193-203).  Replaces
 * train_kernel_gpu<Float, method::lloyd_csr> (train_kernel_lloyd_csr_dpc.cpp:76-242)
 * and the lloyd_csr inference (infer_kernel_dpc.cpp); loop, stop rule, centroid
 * sums and empty-cluster handling are those of km_lloyd_train_f32, with candidate
 * rows copied as in copy_candidates_from_data (empty_cluster_handling_dpc.cpp:152-189). */
#define KM_SPARSE_ZERO_BASED 0
#define KM_SPARSE_ONE_BASED 1
km_status km_lloyd_train_csr_f32(km_ctx* ctx, const float* values, const int64_t* col_indices,
                                 const int64_t* row_offsets, int64_t n, int64_t d, int64_t nnz,
                                 int32_t indexing, const float* C0, int64_t k, int64_t max_iter,
                                 double acc_thr, uint32_t flags, float* C_out, int32_t* labels_out,
                                 double* obj_out, int64_t* iters_out);
km_status km_lloyd_train_csr_f64(km_ctx* ctx, const double* values, const int64_t* col_indices,
                                 const int64_t* row_offsets, int64_t n, int64_t d, int64_t nnz,
                                 int32_t indexing, const double* C0, int64_t k, int64_t max_iter,
                                 double acc_thr, uint32_t flags, double* C_out,
                                 int32_t* labels_out, double* obj_out, int64_t* iters_out);
km_status km_lloyd_infer_csr_f32(km_ctx* ctx, const float* values, const int64_t* col_indices,
                                 const int64_t* row_offsets, int64_t n, int64_t d, int64_t nnz,
                                 int32_t indexing, const float* C, int64_t k, uint32_t flags,
                                 int32_t* labels_out, double* obj_out);
km_status km_lloyd_infer_csr_f64(km_ctx* ctx, const double* values, const int64_t* col_indices,
                                 const int64_t* row_offsets, int64_t n, int64_t d, int64_t nnz,
                                 int32_t indexing, const double* C, int64_t k, uint32_t flags,
                                 int32_t* labels_out, double* obj_out);

/* Initial centroids (kmeans_init).  Methods: */
#define KM_INIT_DENSE 0               /* the first k rows */
#define KM_INIT_RANDOM_DENSE 1        /* k distinct rows, uniformly at random */
#define KM_INIT_PLUS_PLUS_DENSE 2     /* k-means++: first row uniform, then D^2 sampling */
#define KM_INIT_PARALLEL_PLUS_DENSE 3 /* not on the GPU (KM_UNIMPLEMENTED, as the reference) */

/* C_out (k x d) = initial centroids of X (n x d) by `method`, drawn with the
 * counter-based generator keyed by `seed` (plus_plus_dense: D^2 = FP64
 * exact-order distances to the nearest chosen centroid, sampled in a fixed
 * summation order; the result does not depend on timing).  Preconditions and
 * messages of kmeans_init (detail/compute_ops.hpp:40-56): data non-empty,
 * 0 < k < n.  Replaces kmeans_init's GPU compute kernels
 * (backend/gpu/compute_kernels_impl.hpp:36-68 dense,
 * compute_kernel_distr_random_dense_dpc.cpp:33-66 random_dense); the reference
 * runs plus_plus_dense on the host only (train_kernel_lloyd_dense_dpc.cpp:44-74
 * calls DAAL's k-means++ when train has no initial centroids).  Random methods
 * match the reference in distribution, not draw for draw (its engine is DAAL's
 * mt19937). */
km_status km_init_centroids_f32(km_ctx* ctx, const float* X, int64_t n, int64_t d, int64_t k,
                                int32_t method, uint64_t seed, uint32_t flags, float* C_out);
km_status km_init_centroids_f64(km_ctx* ctx, const double* X, int64_t n, int64_t d, int64_t k,
                                int32_t method, uint64_t seed, uint32_t flags, double* C_out);

/* C_out (k x d) = initial centroids of CSR rows (arrays as for km_lloyd_train_csr_*)
 * by a KM_INIT_* method applied to the rows' dense expansion, without forming it:
 * the result equals km_init_centroids_* on the expanded matrix bit for bit (a
 * repeated column holds its values added in stored order).  Replaces the start
 * of a lloyd_csr training without initial centroids, which the reference runs
 * with DAAL's k-means++ on the host (train_kernel_lloyd_csr_dpc.cpp:41-68,
 * kmeans/detail/train_init_centroids.hpp); preconditions and messages as
 * km_init_centroids_*, plus the CSR structure checks. */
km_status km_init_centroids_csr_f32(km_ctx* ctx, const float* values, const int64_t* col_indices,
                                    const int64_t* row_offsets, int64_t n, int64_t d, int64_t nnz,
                                    int32_t indexing, int64_t k, int32_t method, uint64_t seed,
                                    uint32_t flags, float* C_out);
km_status km_init_centroids_csr_f64(km_ctx* ctx, const double* values, const int64_t* col_indices,
                                    const int64_t* row_offsets, int64_t n, int64_t d, int64_t nnz,
                                    int32_t indexing, int64_t k, int32_t method, uint64_t seed,
                                    uint32_t flags, double* C_out);

/* One instance of a batched Lloyd training; all pointers are device memory. */
typedef struct {
    const float* X;      /* n x d row-major points */
    const float* C0;     /* k x d row-major initial centroids */
    float* C_out;        /* k x d final centroids */
    int32_t* labels_out; /* n final labels */
    int64_t n, d, k;
} km_batch_instance_f32;

/* Batched Lloyd training: `count` independent instances (each its own n, d, k),
 * max_iter iterations each (accuracy threshold 0: the reference GPU driver then
 * runs every iteration, train_kernel_lloyd_dense_dpc.cpp:147-168).  Each
 * instance's centroids, labels and objective equal a km_lloyd_train_f32 call on
 * that instance alone, bit for bit.  Every stage of an iteration is one launch over
 * all instances and the iteration needs no host synchronisation; with
 * KM_FLAG_GRAPH the iteration is captured once as a HIP graph and replayed.
 * obj_out and iters_out are host arrays of `count` entries.  The total row count
 * of a batch must stay below 2^31.  No counterpart in the reference, which trains
 * one instance per call (train_kernel_gpu<F, lloyd_dense>::operator(),
 * backend/gpu/train_kernel.hpp:24-29); this is the batched form of that call. */
km_status km_batch_train_f32(km_ctx* ctx, const km_batch_instance_f32* instances, int64_t count,
                             int64_t max_iter, uint32_t flags, double* obj_out,
                             int64_t* iters_out);

/* Autotuner of the batched assignment's tiles: times every tile variant on
 * synthetic instances of each (d, k) bucket's largest shape on this context's
 * device and keeps the fastest per bucket for the context's later batched calls.
 * table_out (may be NULL) receives min(table_len, KM_BATCH_TILE_BUCKETS) chosen
 * variants.  Takes a few seconds and about 1 GiB of device memory while it runs.
 * No counterpart in the reference, whose kernels take fixed launch shapes
 * (kernels_fp_impl.hpp:90-125 sizes the row blocks only). */
km_status km_batch_autotune(km_ctx* ctx, int32_t* table_out, int64_t table_len);

/* One step of a continuous batch, for schedulers that change the set of
 * instances between iterations.  First `iterations` Lloyd iterations of every
 * instance of `iterated` (centroids read from C0, written to C_out, which may
 * alias C0; labels_out unused), then a final pass over every instance of
 * `finished`: labels against C0 into labels_out, objective into obj_out[i]
 * (host array of n_finished entries) and, where C_out is not NULL, the centroids
 * into C_out.  The final pass runs after the iterations, so an instance may be in
 * both sets, its final pass reading the centroids its last iteration wrote.
 * Iterated instances follow the training preconditions, finished ones the
 * inference preconditions; every instance's result equals the single-instance
 * calls bit for bit.  All pointers except obj_out are device memory; flags must be
 * 0; returns after the step has completed.  No counterpart in the reference: it is
 * cluster_updater<F>::update (cluster_updater.hpp:82-179) applied to many
 * instances at once, plus the final assignment of train_kernel_lloyd_dense_dpc.cpp:
 * 169-197. */
km_status km_batch_step_f32(km_ctx* ctx, const km_batch_instance_f32* iterated, int64_t n_iterated,
                            int64_t iterations, const km_batch_instance_f32* finished,
                            int64_t n_finished, uint32_t flags, double* obj_out);

/* ---- Serving --------------------------------------------------------------
 * A request loop over one device.  Fit jobs (fixed iteration count, accuracy
 * threshold 0) and infer jobs are submitted from any thread and return a job id
 * at once.  Jobs are admitted in submission order while the sum of the admitted
 * jobs' HBM footprints stays within the budget; host-memory jobs are then staged
 * into device memory the server owns.  Admitted jobs join the batched lane at its
 * next iteration (continuous batching: each step iterates every active fit job,
 * and jobs finish and leave between steps), except jobs with n*k*d at or above
 * large_job_work, which run one at a time on a second stream through the
 * single-instance engine (BF16 matrix cores when the job asks for
 * KM_FLAG_BF16_REFINE), overlapping the batched lane on the device.  Every job's
 * labels, centroids and objective equal km_lloyd_train_f32 / km_lloyd_infer_f32
 * on that job alone, bit for bit, whatever else runs.  No counterpart in the
 * reference, which serves one train()/infer() call at a time. */
typedef struct km_server km_server;

typedef struct {
    int32_t device;           /* HIP device */
    int32_t max_batch_jobs;   /* active jobs of the batched lane; 0: 4096, at most 65536 */
    int64_t hbm_budget_bytes; /* admission budget over the admitted jobs' footprints */
    int64_t large_job_work;   /* n*k*d from which a job takes the large-job lane;
                                 0: 2^34, negative: never */
    int32_t autotune;         /* nonzero: run km_batch_autotune for the batched lane at creation */
    int32_t reserved;
} km_server_config;

typedef struct {
    int64_t iterations;  /* Lloyd iterations run (fit), 0 (infer) */
    double objective;    /* sum of squared distances to the final centroids */
    double queue_ms;     /* submission -> admission */
    double latency_ms;   /* submission -> results in place */
    int32_t lane;        /* 0 batched lane, 1 large-job lane */
    int32_t reserved;
} km_job_result;

typedef struct {
    int64_t submitted, completed, failed;
    int64_t queued;             /* jobs waiting for admission now */
    int64_t steps;              /* batched-lane steps run */
    int64_t step_jobs;          /* jobs summed over those steps */
    int64_t max_step_jobs;      /* most jobs in one step */
    int64_t large_jobs;         /* jobs run on the large-job lane */
    int64_t bytes_admitted;     /* footprints of the admitted, unfinished jobs now */
    int64_t peak_bytes_admitted;
} km_server_stats;

km_status km_server_create(const km_server_config* cfg, km_server** out);
/* Runs every submitted job to completion, then frees the server.  No other
 * call on the server may overlap it or follow it. */
km_status km_server_destroy(km_server* server);
/* Message of this thread's last failed km_server_* call (or failed job's wait). */
const char* km_server_last_error(void);

/* Fit job: max_iter Lloyd iterations from C0.  With KM_FLAG_DEVICE_POINTERS, X,
 * C0, C_out and labels_out are device memory used in place; otherwise host
 * memory, copied in at admission and out at completion.  Every buffer must stay
 * valid until the job has finished.  Preconditions as km_lloyd_train_f32; a job
 * whose footprint exceeds the budget fails with KM_DEVICE_BAD_ALLOC. */
km_status km_server_submit_fit_f32(km_server* server, const float* X, int64_t n, int64_t d,
                                   const float* C0, int64_t k, int64_t max_iter, uint32_t flags,
                                   float* C_out, int32_t* labels_out, int64_t* job_out);
/* Infer job: labels of X against the centroids C (k x d) and the objective. */
km_status km_server_submit_infer_f32(km_server* server, const float* X, int64_t n, int64_t d,
                                     const float* C, int64_t k, uint32_t flags, int32_t* labels_out,
                                     int64_t* job_out);
/* Wait up to timeout_ms (negative: no limit) for a job.  Returns KM_PENDING if it
 * is still running; otherwise the job's own status, fills *out and forgets the job. */
km_status km_server_wait(km_server* server, int64_t job, int64_t timeout_ms, km_job_result* out);
km_status km_server_get_stats(km_server* server, km_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* ONEDAL_AMD_KMEANS_HIP_H */
