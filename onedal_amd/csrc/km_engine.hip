// km_engine.hip -- host driver and C ABI of the MI355X k-means Lloyd engine.
//
// Replaces train_kernel_gpu<F, lloyd_dense> (cpp/oneapi/dal/algo/kmeans/backend/
// gpu/train_kernel_lloyd_dense_dpc.cpp:76-207) with its cluster_updater
// (cluster_updater.hpp:82-179) and infer_kernel_gpu (infer_kernel_dpc.cpp:36-107).
//
// One Lloyd iteration on the context's stream:
//   centroid_prep -> assign (+ exact re-score) -> count -> scan -> stable sort of
//   row ids by label -> chunk sums (FP64) -> merge/divide [-> objective]
//   -> one small D2H (empty-cluster count, objective) and a stream sync
//   [-> empty-cluster refill].
// The reference syncs twice per iteration (cluster_updater.hpp:135-136 and
// kernels_integral_dpc.cpp:148); this driver syncs once.

#include <hip/hip_runtime.h>

#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <new>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/onedal_amd/kmeans_hip.h"
#include "km_args.hpp"
#include "km_assign.hpp"
#include "km_assign_bf16.hpp"
#include "km_batch.hpp"
#include "km_common.hpp"
#include "km_csr.hpp"
#include "km_init.hpp"
#include "km_probe.hpp"
#include "km_rng.hpp"
#include "km_refine.hpp"
#include "km_update.hpp"

namespace km {

#define KM_HIP(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            throw km::status_error(e_ == hipErrorOutOfMemory ? KM_DEVICE_BAD_ALLOC             \
                                                         : KM_INTERNAL_ERROR,              \
                               std::string(#expr) + ": " + hipGetErrorString(e_));         \
        }                                                                                  \
    } while (0)

#define KM_LAUNCH_CHECK() KM_HIP(hipGetLastError())

// size of the fixed scratch array of km_lloyd_infer_f32_noalloc
constexpr size_t NOALLOC_BYTES = size_t(64) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

// grow-only device buffer
struct dbuf {
    void* p = nullptr;
    size_t bytes = 0;
    template <typename T>
    T* get(size_t count) {
        const size_t need = std::max<size_t>(count * sizeof(T), 256);
        if (need > bytes) {
            if (p) (void)hipFree(p);
            p = nullptr;
            bytes = 0;
            hipError_t e = hipMalloc(&p, need);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                throw status_error(KM_DEVICE_BAD_ALLOC,
                                   "device allocation of " + std::to_string(need) + " bytes failed");
            }
            bytes = need;
        }
        return static_cast<T*>(p);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

// workspace of one batched plan: instance and prefix tables, concatenated
// per-row, per-cluster and per-chunk arrays, sort scratch
struct batch_ws {
    dbuf inst, pre, labels, keys, perm, d64, ccur, cnext, counts, offsets, cb, empty, meta, qpre,
        part, objc, obj, sort_tmp, asel, apre, bhist, ticket;
    void release() {
        dbuf* all[] = {&inst, &pre,  &labels, &keys, &perm,     &d64,  &ccur, &cnext, &counts, &offsets, &cb,
                       &empty, &meta, &qpre,  &part, &objc, &obj, &sort_tmp, &asel, &apre, &bhist, &ticket};
        for (dbuf* b : all) b->release();
    }
};

}  // namespace km

// Workspace of km_lloyd_infer_f32_noalloc: one copy per device, held until the
// process ends.  In host code this name is a stand-in, not a device address;
// the address comes from hipGetSymbolAddress (km::scratch_address).
static __device__ __align__(256) unsigned char scratch[km::NOALLOC_BYTES];

struct km_ctx {
    int device = 0;
    int cu_count = 256;
    hipStream_t stream = nullptr;
    std::string last_error;
    km_stats stats{};
    int assign_path_opt = 0;
    bool timing = false;
    // batched assignment tile variant per (d, k) bucket (-1: default by k), and a
    // forced variant (0: none, v + 1: variant v)
    int8_t tile_table[KM_BATCH_TILE_BUCKETS];
    int tile_force = 0;
    km_ctx() { std::memset(tile_table, -1, sizeof(tile_table)); }

    // workspace
    km::dbuf x_bf16, enorm, xmax, c_bf16, cnb, params, coll_rows, coll_T, coll_count, coll_ids;
    km::dbuf seg_cand_rows, seg_cand_cnt, seg_cand_ids, seg_coll_rows, seg_coll_T, seg_counts;
    km::dbuf x_dev, x_pad, c_cur, c_next, cn, cpad, cmax, xnorm, labels, lists_rows, lists_cnt, lists_ids,
        lists_full, lists_counters, counts, offsets, chunk_base, empty_list, meta, perm, keys_out,
        sort_tmp, part, d64, obj_chunks, obj_dev, hist, cand_gt, cand_eq, cand_dev,
        c_out_dev;
    // batched training and steps: the iterated set, and the finished set of a step
    km::batch_ws bws, bws_fin;
    km::dbuf tune_x, tune_c, tune_l;  // autotuner inputs (released after tuning)
    km::dbuf init_d2, init_chunks, init_chosen, init_ids;  // initialisation
    km::dbuf gt_d;  // empty-cluster candidates' d64
    km::dbuf stat_acc, coop_wcnt;  // device-side statistics; per-workgroup counts of ec_count_kernel
    km::dbuf cs_inst, cs_pre, cs_hist, cs_gsum;  // single-instance counting sort (label_sort)
    // CSR path: values, column indices, row offsets (when uploaded), row squares,
    // transposed centroids, centroid squares, closest distances, structure flag
    km::dbuf csr_vals, csr_cols, csr_offs, csr_xs, csr_ct, csr_cs, csr_closest, csr_flag;
    int32_t* pinned = nullptr;  // small pinned host staging: meta[2], counters[2], obj (as 2 ints)
    // this device's copy of the scratch array (null if its address is unknown)
    void* scratch = nullptr;
    std::string scratch_err;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;

    hipEvent_t event() {
        if (ev_used == ev_pool.size()) {
            hipEvent_t e;
            KM_HIP(hipEventCreate(&e));
            ev_pool.push_back(e);
        }
        return ev_pool[ev_used++];
    }
};

namespace km {

static int grid_for(const km_ctx* ctx, int64_t work_items, int per_block, int waves_per_cu = 8) {
    const int64_t want = (work_items + per_block - 1) / per_block;
    const int64_t cap = int64_t(ctx->cu_count) * waves_per_cu;
    return int(std::max<int64_t>(1, std::min(want, cap)));
}

template <typename T>
static void fill(km_ctx* ctx, T* p, int64_t count, T value) {
    const int grid = grid_for(ctx, count, 256, 4);
    hipLaunchKernelGGL(fill_kernel<T>, dim3(grid), dim3(256), 0, ctx->stream, p, count, value);
    KM_LAUNCH_CHECK();
}

// Fixed device memory handed out front to back (256-byte aligned pieces); never
// allocates.  Running past the end is an error, not a fault.
struct arena {
    unsigned char* base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    template <typename T>
    T* get(size_t count) {
        const size_t off = (used + 255) & ~size_t(255);
        const size_t bytes = std::max<size_t>(count * sizeof(T), 1);
        if (off > cap || bytes > cap - off)
            throw status_error(KM_INTERNAL_ERROR, "workspace plan exceeds the fixed scratch buffer");
        used = off + bytes;
        return reinterpret_cast<T*>(base + off);
    }
};

// workspace piece: from the arena when one is given, else the context's buffer
template <typename T>
static T* take(km_ctx* ctx, dbuf km_ctx::*buf, arena* A, size_t count) {
    return A ? A->get<T>(count) : (ctx->*buf).get<T>(count);
}

template <typename F>
struct work {
    // sizes
    int64_t n = 0;
    int d = 0, k = 0, dpad = 0;
    // device views
    const F* X = nullptr;
    const F* Xv = nullptr;    // rows with stride ldx (multiple of 4, 16-byte aligned) for float4 reads
    int64_t ldx = 0;
    float* xnorm = nullptr;
    int32_t* labels = nullptr;
    refine_lists L{};
    F* cn = nullptr;
    F* cpad = nullptr;
    unsigned long long* cmax_bits = nullptr;  // [0] max ||c|| ([1], [2]: BF16 path)
    // BF16 path
    uint16_t* Xb = nullptr;       // n x dpad bf16 rows
    float* enorm = nullptr;       // ||x - bf16(x)|| (up)
    unsigned long long* xmax_bits = nullptr;
    uint16_t* Cb = nullptr;       // k x dpad bf16(-2 c)
    float* cnb = nullptr;         // ||c||^2 + B (rounded up), 32 * ceil(k / 32) entries
    bf16_params* params = nullptr;
    // workspace source: a fixed arena, or (null) the context's buffers
    arena* A = nullptr;
};

static int pad_dim(int d) {
    int p = 32;
    while (p < d) p <<= 1;
    return p;
}

// 1 generic tile kernel, 2 matrix cores of the data's precision (FP32 or FP64),
// 3 BF16 matrix cores + exact re-score.  BF16 applies to float data with
// d <= 256; double data takes path 2 for d <= 256, else path 1.
template <typename F>
static int choose_path(const km_ctx* ctx, int d, int k, uint32_t flags) {
    if (std::is_same<F, double>::value) return (d <= 256 && ctx->assign_path_opt != 1) ? 2 : 1;
    if (d > 256) return 1;
    // keys carry (block, register) in at most 15 low bits: k <= 65536
    if ((flags & KM_FLAG_BF16_REFINE) && k <= 65536) return 3;
    if (ctx->assign_path_opt == 1) return 1;
    return 2;
}

template <typename F>
static band_consts make_band(int d) {
    const double u = fp_traits<F>::u;
    const double ua = fp_traits<F>::u_acc;
    band_consts b;
    const double ga = gamma_m(double(d), ua);
    b.a_xc = 2.0 * ga + 2.01 * u;
    b.a_cc = 2.01 * u + gamma_m(double(d), std::is_same<F, double>::value ? ua : 2.220446049250313e-16);
    const double g = gamma_m(double(d + 2), u);
    b.g_canon = g / (1.0 - g);
    return b;
}

static bf16_consts make_bf16_consts(int d, int k) {
    bf16_consts c;
    int mbits = 0;
    while ((1 << mbits) < bf16_padded_k(k) / 32) mbits++;
    c.ib = 4 + mbits;
    c.u = fp_traits<float>::u;
    c.g_acc = gamma_m(double(d + 1), fp_traits<float>::u_acc);
    c.trunc = std::ldexp(1.0, c.ib - 23);
    const double g = gamma_m(double(d + 2), fp_traits<float>::u);
    c.g_canon = g / (1.0 - g);
    return c;
}

// per-centroid prep: ||c||^2, padded copy, max ||c||
template <typename F>
static void centroid_prep(km_ctx* ctx, work<F>& w, const F* C, int path) {
    if constexpr (std::is_same<F, float>::value) {
        if (path == 3) {
            fill<unsigned long long>(ctx, w.cmax_bits, 3, 0ull);
            const int kp = bf16_padded_k(w.k);
            // four centroids per wave, so few workgroups add to the three maxima
            const int grid = grid_for(ctx, int64_t(kp) * 16, 256, 4);
            hipLaunchKernelGGL(centroid_prep_bf16_kernel, dim3(grid), dim3(256), 0, ctx->stream, C,
                               w.k, kp, w.d, w.dpad, w.Cb, w.cn, w.cmax_bits);
            KM_LAUNCH_CHECK();
            hipLaunchKernelGGL(bf16_params_kernel, dim3(1), dim3(64), 0, ctx->stream, w.cmax_bits,
                               w.xmax_bits, make_bf16_consts(w.d, w.k), w.params);
            KM_LAUNCH_CHECK();
            hipLaunchKernelGGL(cnb_kernel, dim3(grid_for(ctx, kp, 256, 1)), dim3(256), 0, ctx->stream,
                               w.cn, w.k, kp, w.params, w.cnb);
            KM_LAUNCH_CHECK();
            return;
        }
    }
    const bool need_pad = path == 2;
    fill<unsigned long long>(ctx, w.cmax_bits, 1, 0ull);
    const int grid = grid_for(ctx, int64_t(w.k) * 64, 256, 4);
    hipLaunchKernelGGL(centroid_prep_kernel<F>, dim3(grid), dim3(256), 0, ctx->stream, C, w.k, w.d,
                       w.dpad, w.cn, need_pad ? w.cpad : (F*)nullptr, w.cmax_bits);
    KM_LAUNCH_CHECK();
}

template <int DP, int NB, bool RES>
static void launch_mfma_f32_mode(km_ctx* ctx, work<float>& w, const band_consts& bc, int grid,
                                 size_t lds) {
    auto kern = assign_mfma_f32_kernel<DP, NB, RES>;
    KM_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), lds, ctx->stream, w.Xv, w.ldx, w.n, w.d, w.cpad,
                       w.k, w.cn, w.xnorm, (const double*)w.cmax_bits, bc, w.labels, w.L);
    KM_LAUNCH_CHECK();
}

template <int DP, int NB>
static void launch_mfma_f32(km_ctx* ctx, work<float>& w, const band_consts& bc) {
    constexpr int ROWS_PER_WG = 4 * 32 * NB;
    const int64_t tiles = (w.n + ROWS_PER_WG - 1) / ROWS_PER_WG;
    const int grid = int(std::max<int64_t>(1, std::min<int64_t>(tiles, int64_t(ctx->cu_count) * 2)));
    // centroids resident in LDS when two workgroups still fit a CU (160 KB)
    const size_t res_lds = mfma_f32_lds_bytes(DP, w.k, true);
    if (res_lds <= 78 * 1024)
        launch_mfma_f32_mode<DP, NB, true>(ctx, w, bc, grid, res_lds);
    else
        launch_mfma_f32_mode<DP, NB, false>(ctx, w, bc, grid, mfma_f32_lds_bytes(DP, w.k, false));
}

template <int DP, int NB>
static void launch_mfma_f64(km_ctx* ctx, work<double>& w, const band_consts& bc) {
    constexpr int ROWS_PER_WG = 4 * 16 * NB;
    const int64_t tiles = (w.n + ROWS_PER_WG - 1) / ROWS_PER_WG;
    const int grid = int(std::max<int64_t>(1, std::min<int64_t>(tiles, int64_t(ctx->cu_count) * 2)));
    const size_t lds = mfma_f64_lds_bytes(DP);
    auto kern = assign_mfma_f64_kernel<DP, NB>;
    KM_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), lds, ctx->stream, w.Xv, w.ldx, w.n, w.d, w.cpad, w.k, w.cn,
                       w.xnorm, (const double*)w.cmax_bits, bc, w.labels, w.L);
    KM_LAUNCH_CHECK();
}

template <int DPB>
static void launch_bf16(km_ctx* ctx, work<float>& w, bool collect) {
    constexpr int NB = DPB >= 256 ? 1 : 2;  // row registers of two blocks fit below 256 VGPRs
    constexpr int ROWS_PER_WG = 4 * 32 * NB;
    const size_t lds = bf16_lds_bytes(DPB);
    const bf16_consts kc = make_bf16_consts(w.d, w.k);
    if (!collect) {
        const int64_t tiles = (w.n + ROWS_PER_WG - 1) / ROWS_PER_WG;
        const int grid = int(std::max<int64_t>(1, std::min<int64_t>(tiles, int64_t(ctx->cu_count) * 2)));
        // one list segment per wave, as long as the rows that wave can see
        const int nseg = grid * 4;
        seg_lists S;
        S.cap = (tiles + grid - 1) / grid * (32 * NB);
        const int64_t ents = S.cap * nseg;
        S.cand_rows = take<int32_t>(ctx, &km_ctx::seg_cand_rows, w.A, ents);
        S.cand_count = take<int32_t>(ctx, &km_ctx::seg_cand_cnt, w.A, ents);
        S.cand_ids = take<int32_t>(ctx, &km_ctx::seg_cand_ids, w.A, ents * MAXC);
        S.coll_rows = take<int32_t>(ctx, &km_ctx::seg_coll_rows, w.A, ents);
        S.coll_T = take<double>(ctx, &km_ctx::seg_coll_T, w.A, ents);
        S.counts = take<int32_t>(ctx, &km_ctx::seg_counts, w.A, 2 * int64_t(nseg));
        auto kern = assign_bf16_kernel<DPB, NB, false>;
        KM_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   int(lds)));
        hipLaunchKernelGGL(kern, dim3(grid), dim3(256), lds, ctx->stream, w.Xb, w.n, w.Cb, w.k, w.cnb,
                           w.xnorm, w.enorm, w.params, kc, w.labels, w.L, S);
        KM_LAUNCH_CHECK();
        hipLaunchKernelGGL(compact_lists_kernel, dim3((nseg + 3) / 4), dim3(256), 0, ctx->stream, S,
                           nseg, w.L);
    } else {
        // row count is on the device; a fixed grid strides over it.  One row block
        // per wave (128-row tiles, half the hit-list LDS) and three workgroups per
        // CU: the few collect rows spread over more, shorter tiles (0.28 ms per
        // pass at n=2e7, d=128, k=1024 against 0.36 ms with 256-row tiles, two per CU)
        const size_t clds = bf16_collect_lds_bytes(DPB, 4 * 32);
        auto ckern = assign_bf16_kernel<DPB, 1, true>;
        KM_HIP(hipFuncSetAttribute((const void*)ckern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   int(clds)));
        hipLaunchKernelGGL(ckern, dim3(ctx->cu_count * 3), dim3(256), clds, ctx->stream, w.Xb, w.n,
                           w.Cb, w.k, w.cnb, w.xnorm, w.enorm, w.params, kc, w.labels, w.L,
                           seg_lists{});
    }
    KM_LAUNCH_CHECK();
}

// exact re-score of a list of (row, candidate ids) entries; overflow -> full list
template <typename F>
static void refine_pairs(km_ctx* ctx, work<F>& w, const F* C, const int32_t* rows,
                         const int32_t* counts, const int32_t* ids, int stride,
                         const int32_t* count_p) {
    const int64_t row_bytes = int64_t(refine_row_stride<F>(w.d)) * sizeof(F);
    const int rb = int(std::max<int64_t>(1, std::min<int64_t>(16, 16384 / row_bytes)));
    const size_t lds = size_t(4) * rb * row_bytes;
    auto kern = refine_pairs_kernel<F>;
    KM_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    hipLaunchKernelGGL(kern, dim3(ctx->cu_count * 4), dim3(256), lds, ctx->stream, w.X, w.d, C, rows,
                       counts, ids, stride, count_p, rb, w.L.full_rows, w.L.counters + 1, w.labels);
    KM_LAUNCH_CHECK();
}

// assignment of every row against w.X/C (centroid_prep already done for C)
template <typename F>
static void assign(km_ctx* ctx, work<F>& w, const F* C, int path) {
    fill<int32_t>(ctx, w.L.counters, 3, 0);
    const band_consts bc = make_band<F>(w.d);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->timing) {
        e0 = ctx->event();
        e1 = ctx->event();
        KM_HIP(hipEventRecord(e0, ctx->stream));
    }
    if constexpr (std::is_same<F, float>::value) {
        if (path == 2) {
            switch (w.dpad) {
                case 32: launch_mfma_f32<32, 2>(ctx, w, bc); break;
                case 64: launch_mfma_f32<64, 2>(ctx, w, bc); break;
                case 128: launch_mfma_f32<128, 1>(ctx, w, bc); break;
                default: launch_mfma_f32<256, 1>(ctx, w, bc); break;
            }
        }
        if (path == 3) {
            switch (w.dpad) {
                case 32: launch_bf16<32>(ctx, w, false); break;
                case 64: launch_bf16<64>(ctx, w, false); break;
                case 128: launch_bf16<128>(ctx, w, false); break;
                default: launch_bf16<256>(ctx, w, false); break;
            }
        }
    }
    if constexpr (std::is_same<F, double>::value) {
        if (path == 2) {
            switch (w.dpad) {
                case 32: launch_mfma_f64<32, 2>(ctx, w, bc); break;
                case 64: launch_mfma_f64<64, 2>(ctx, w, bc); break;
                case 128: launch_mfma_f64<128, 2>(ctx, w, bc); break;
                default: launch_mfma_f64<256, 1>(ctx, w, bc); break;
            }
        }
    }
    if (path == 1) {
        const int64_t tiles = (w.n + 63) / 64;
        const int grid = int(std::max<int64_t>(1, std::min<int64_t>(tiles, int64_t(ctx->cu_count) * 4)));
        hipLaunchKernelGGL(assign_generic_kernel<F>, dim3(grid), dim3(256), 0, ctx->stream, w.X, w.n,
                           w.d, C, w.k, w.cn, w.xnorm, (const double*)w.cmax_bits, bc, w.labels, w.L);
        KM_LAUNCH_CHECK();
    }
    if (ctx->timing) {
        KM_HIP(hipEventRecord(e1, ctx->stream));
    }
    ctx->stats.assign_launches++;
    const int rgrid = ctx->cu_count * 4;
    refine_pairs<F>(ctx, w, C, w.L.cand_rows, w.L.cand_count, w.L.cand_ids, MAXC, w.L.counters + 0);
    if constexpr (std::is_same<F, float>::value) {
        if (path == 3) {
            switch (w.dpad) {
                case 32: launch_bf16<32>(ctx, w, true); break;
                case 64: launch_bf16<64>(ctx, w, true); break;
                case 128: launch_bf16<128>(ctx, w, true); break;
                default: launch_bf16<256>(ctx, w, true); break;
            }
            refine_pairs<F>(ctx, w, C, w.L.coll_rows, w.L.coll_count, w.L.coll_ids, MAXC2,
                            w.L.counters + 2);
        }
    }
    if (w.d % 4 == 0 && reinterpret_cast<uintptr_t>(w.X) % (4 * sizeof(F)) == 0 &&
        reinterpret_cast<uintptr_t>(C) % (4 * sizeof(F)) == 0)
        hipLaunchKernelGGL(refine_full_vec_kernel<F>, dim3(rgrid), dim3(256), 0, ctx->stream, w.X,
                           w.d, C, w.k, w.L, w.labels);
    else
        hipLaunchKernelGGL(refine_full_kernel<F>, dim3(rgrid), dim3(256), 0, ctx->stream, w.X, w.d,
                           C, w.k, w.L, w.labels);
    KM_LAUNCH_CHECK();
}

// objective chunk sums (objective_chunks_kernel)
template <typename F>
static void launch_objective(km_ctx* ctx, int grid, const F* X, int64_t n, int d, const F* C,
                             const int32_t* labels, double* d64_out, double* chunks,
                             const int32_t* gate = nullptr) {
    const size_t lds = objective_lds_bytes<F>();
    hipLaunchKernelGGL(objective_chunks_kernel<F>, dim3(grid), dim3(256), lds, ctx->stream, X, n, d, C, labels,
                       d64_out, chunks, gate);
    KM_LAUNCH_CHECK();
}

// objective of the current labels against C; optional per-row d64 output
template <typename F>
static double* objective(km_ctx* ctx, work<F>& w, const F* C, double* d64_out) {
    const int64_t nchunks = (w.n + OBJ_CHUNK - 1) / OBJ_CHUNK;
    double* chunks = ctx->obj_chunks.get<double>(nchunks);
    double* out = ctx->obj_dev.get<double>(1);
    const int grid = int(std::min<int64_t>(nchunks, int64_t(ctx->cu_count) * 8));
    launch_objective<F>(ctx, grid, w.X, w.n, w.d, C, w.labels, d64_out, chunks);
    KM_LAUNCH_CHECK();
    hipLaunchKernelGGL(sum_sequential_kernel, dim3(1), dim3(64), 0, ctx->stream, chunks, nchunks, out);
    KM_LAUNCH_CHECK();
    return out;
}

template <typename F>
static void launch_reduce(km_ctx* ctx, work<F>& w, int32_t* perm, int64_t* offsets,
                          int32_t* chunk_base, int32_t* meta, double* part) {
    const int64_t max_chunks = w.n / REDUCE_CHUNK + w.k + 1;
    if (w.d > 32 && w.d <= 512 && w.d % 4 == 0 &&
        reinterpret_cast<uintptr_t>(w.X) % (4 * sizeof(F)) == 0) {
        // sub-wave per chunk with 4-wide loads: lpc lanes cover the row V times
        const int V = w.d <= 128 ? 1 : (w.d <= 256 ? 2 : 4);
        int lpc = 1;
        while (lpc * 4 * V < w.d) lpc <<= 1;
        const int subs = 64 / lpc;
        const int grid = grid_for(ctx, (max_chunks + subs - 1) / subs * 64, 256, 8);
#define KM_REDV(V_)                                                                                 \
    hipLaunchKernelGGL((reduce_chunks_vec_kernel<F, V_>), dim3(grid), dim3(256), 0, ctx->stream,    \
                       w.X, w.d, w.k, perm, offsets, chunk_base, meta, part, lpc)
        if (V == 1) KM_REDV(1);
        else if (V == 2) KM_REDV(2);
        else KM_REDV(4);
#undef KM_REDV
        KM_LAUNCH_CHECK();
        return;
    }
    const int grid = grid_for(ctx, max_chunks * 64, 256, 8);
    const int cpl = w.d <= 64 ? 1 : (w.d + 63) / 64;
#define KM_RED(C_)                                                                                  \
    hipLaunchKernelGGL((reduce_chunks_kernel<F, C_>), dim3(grid), dim3(256), 0, ctx->stream, w.X,   \
                       w.d, w.k, perm, offsets, chunk_base, meta, part)
    if (cpl <= 1) KM_RED(1);
    else if (cpl <= 2) KM_RED(2);
    else if (cpl <= 4) KM_RED(4);
    else if (cpl <= 8) KM_RED(8);
    else if (cpl <= 16) KM_RED(16);
    else if (cpl <= 32) KM_RED(32);
    else if (cpl <= 64) KM_RED(64);
    else throw status_error(KM_UNIMPLEMENTED, "column count above 4096 is not supported");
#undef KM_RED
    KM_LAUNCH_CHECK();
}

// stable grouping of row ids by label
static void sort_by_label(km_ctx* ctx, const int32_t* labels, int64_t n, int k, int32_t* perm) {
    unsigned bits = 1;
    while ((1ull << bits) < (unsigned long long)k) bits++;
    int32_t* keys_out = ctx->keys_out.get<int32_t>(n);
    rocprim::counting_iterator<int32_t> ids(0);
    size_t tmp_bytes = 0;
    KM_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, labels, keys_out, ids, perm,
                                     (unsigned int)n, 0, bits, ctx->stream));
    void* tmp = ctx->sort_tmp.get<char>(tmp_bytes);
    KM_HIP(rocprim::radix_sort_pairs(tmp, tmp_bytes, labels, keys_out, ids, perm, (unsigned int)n, 0,
                                     bits, ctx->stream));
}

// Label counts and the stable grouping of row ids by label for one instance.
// k <= BATCH_HIST_MAX: the batched path's counting sort on a one-instance table
// (per-block histograms that also add into counts, block destinations from
// sort_group_sums_kernel / sort_block_bases_kernel, the LDS-staged scatter);
// otherwise count_kernel and rocPRIM's radix sort.  Both leave the members of
// each cluster in ascending row order.
struct label_sort {
    bool csort = false;
    int64_t n = 0, nb = 0;
    int k = 0, G = 0, GB = 0;
    batch_inst* I = nullptr;
    int64_t* pre = nullptr;
    int32_t* bhist = nullptr;
    int32_t* gsum = nullptr;
};

// the scatter kernel's LDS limit, set to the largest use (every caller sets the
// same value, so concurrent contexts never lower it under each other)
static void allow_scatter_lds() {
    KM_HIP(hipFuncSetAttribute((const void*)batch_scatter_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(batch_scatter_lds_bytes(BATCH_HIST_MAX))));
}

static label_sort label_sort_setup(km_ctx* ctx, int64_t n, int k) {
    label_sort ls;
    ls.n = n;
    ls.k = k;
    if (k > BATCH_HIST_MAX) return ls;
    ls.csort = true;
    ls.nb = (n + BATCH_CNT_ROWS - 1) / BATCH_CNT_ROWS;
    ls.G = int(std::min<int64_t>(ls.nb, 64));
    ls.GB = int((ls.nb + ls.G - 1) / ls.G);
    ls.I = ctx->cs_inst.get<batch_inst>(1);
    ls.pre = ctx->cs_pre.get<int64_t>(2);
    ls.bhist = ctx->cs_hist.get<int32_t>(ls.nb * k);
    ls.gsum = ctx->cs_gsum.get<int32_t>(int64_t(ls.G) * k);
    batch_inst b{};
    b.n = n;
    b.k = k;
    const int64_t pre[2] = {0, ls.nb};
    KM_HIP(hipMemcpyAsync(ls.I, &b, sizeof(b), hipMemcpyHostToDevice, ctx->stream));
    KM_HIP(hipMemcpyAsync(ls.pre, pre, sizeof(pre), hipMemcpyHostToDevice, ctx->stream));
    KM_HIP(hipStreamSynchronize(ctx->stream));  // b and pre are stack values
    allow_scatter_lds();
    return ls;
}

// counts of the labels (with the counting sort: the block histograms, their
// group sums, and the counts as the column sums of those)
static void label_counts(km_ctx* ctx, const label_sort& ls, const int32_t* labels, int32_t* counts) {
    if (ls.csort) {
        const int grid = int(std::min<int64_t>(ls.nb, int64_t(ctx->cu_count) * 8));
        hipLaunchKernelGGL(batch_count_kernel, dim3(grid), dim3(256), 0, ctx->stream, ls.I, 1, ls.pre, labels,
                           nullptr, ls.bhist);
        KM_LAUNCH_CHECK();
        const int g2 = int((int64_t(ls.G) * ls.k + 255) / 256);
        hipLaunchKernelGGL(sort_group_sums_kernel, dim3(g2), dim3(256), 0, ctx->stream, ls.bhist, ls.nb, ls.k, ls.G,
                           ls.GB, ls.gsum);
        KM_LAUNCH_CHECK();
        hipLaunchKernelGGL(sort_counts_kernel, dim3((ls.k + 255) / 256), dim3(256), 0, ctx->stream, ls.gsum, ls.k,
                           ls.G, counts);
    } else {
        fill<int32_t>(ctx, counts, ls.k, 0);
        const bool lds = ls.k <= 16384;
        hipLaunchKernelGGL(count_kernel, dim3(grid_for(ctx, ls.n, 256, 4)), dim3(256),
                           lds ? ls.k * sizeof(int32_t) : 0, ctx->stream, labels, ls.n, ls.k, counts, lds ? 1 : 0);
    }
    KM_LAUNCH_CHECK();
}

// row ids grouped by label (after the scan has set offsets)
static void label_group(km_ctx* ctx, const label_sort& ls, const int32_t* labels, const int64_t* offsets,
                        int32_t* perm) {
    if (!ls.csort) {
        sort_by_label(ctx, labels, ls.n, ls.k, perm);
        return;
    }
    // the group sums are label_counts' (bhist is unchanged since)
    const int g2 = int((int64_t(ls.G) * ls.k + 255) / 256);
    hipLaunchKernelGGL(sort_block_bases_kernel, dim3(g2), dim3(256), 0, ctx->stream, ls.bhist, ls.nb, ls.k, ls.G,
                       ls.GB, ls.gsum, offsets);
    KM_LAUNCH_CHECK();
    const int grid = int(std::min<int64_t>(ls.nb, int64_t(ctx->cu_count) * 8));
    hipLaunchKernelGGL(batch_scatter_kernel, dim3(grid), dim3(256), batch_scatter_lds_bytes(ls.k), ctx->stream,
                       ls.I, 1, ls.pre, labels, ls.bhist, perm, ls.k);
    KM_LAUNCH_CHECK();
}

// Empty clusters of one training iteration, on the device (the ec_* kernels: a
// fixed sequence of launches, each ending at once when meta[0] = 0): the m rows
// with the largest d64 (ties: lowest row) go, in that order, to the empty
// clusters in ascending index order (dense rows X; with X null the picks are
// left in the context's candidate list for the sparse fill); the objective at
// objp (when not null) drops by their d64, subtracted in that order.  The
// sequence also adds the re-score counters (when not null) into acc.  Returns
// the candidate list.
template <typename F>
static int32_t* launch_empty_clusters(km_ctx* ctx, const F* X, int64_t n, int d, int k, const int32_t* meta,
                                      const int32_t* empty_list, const int32_t* counters, F* Cnext,
                                      const double* d64, double* objp, unsigned long long* acc) {
    const int grid = int(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, int64_t(ctx->cu_count) * 4)));
    unsigned int* hist = ctx->hist.get<unsigned int>(EC_PASSES * 256);
    int32_t* wcnt = ctx->coop_wcnt.get<int32_t>(2 * int64_t(grid));
    int32_t* gt = ctx->cand_gt.get<int32_t>(k);
    int32_t* eq = ctx->cand_eq.get<int32_t>(k);
    int32_t* cand = ctx->cand_dev.get<int32_t>(k);
    double* cand_d = ctx->gt_d.get<double>(k);
    hipLaunchKernelGGL(ec_begin_kernel, dim3(1), dim3(256), 0, ctx->stream, meta, counters, acc, hist);
    KM_LAUNCH_CHECK();
    for (int pass = 0; pass < EC_PASSES; ++pass) {
        hipLaunchKernelGGL(ec_hist_kernel, dim3(grid), dim3(256), 0, ctx->stream, d64, n, meta, hist, pass);
        KM_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(ec_count_kernel, dim3(grid), dim3(256), 0, ctx->stream, d64, n, meta, hist, wcnt);
    KM_LAUNCH_CHECK();
    hipLaunchKernelGGL(ec_collect_kernel, dim3(grid), dim3(256), 0, ctx->stream, d64, n, meta, hist, wcnt, gt,
                       eq);
    KM_LAUNCH_CHECK();
    hipLaunchKernelGGL(ec_order_kernel, dim3(grid), dim3(256), 0, ctx->stream, d64, meta, hist, gt, eq, cand,
                       cand_d);
    KM_LAUNCH_CHECK();
    hipLaunchKernelGGL(ec_finish_kernel<F>, dim3(grid), dim3(256), 0, ctx->stream, X, d, meta, empty_list, cand,
                       cand_d, Cnext, objp, acc);
    KM_LAUNCH_CHECK();
    return cand;
}

// Dense rows: the d64 pass first (the objective pass, gated on meta[0] > 0
// unless objp already ran with d64 out), then launch_empty_clusters.
template <typename F>
static void empty_clusters(km_ctx* ctx, work<F>& w, const F* Ccur, F* Cnext, const int32_t* empty_list,
                           const int32_t* meta, double* objp, double* d64, unsigned long long* acc) {
    if (!objp) {
        const int64_t nchunks = (w.n + OBJ_CHUNK - 1) / OBJ_CHUNK;
        const int ogrid = int(std::min<int64_t>(nchunks, int64_t(ctx->cu_count) * 8));
        launch_objective<F>(ctx, ogrid, w.X, w.n, w.d, Ccur, w.labels, d64,
                            ctx->obj_chunks.get<double>(nchunks), meta);
    }
    launch_empty_clusters<F>(ctx, w.X, w.n, w.d, w.k, meta, empty_list, w.L.counters, Cnext, d64, objp, acc);
}

template <typename F>
static void setup_work(km_ctx* ctx, work<F>& w, const F* Xd, int64_t n, int d, int k,
                       int path, arena* A = nullptr, int32_t* labels_into = nullptr) {
    const bool vec_view = path == 2;
    w.n = n;
    w.d = d;
    w.k = k;
    w.dpad = pad_dim(d);
    w.X = Xd;
    w.A = A;
    w.xnorm = take<float>(ctx, &km_ctx::xnorm, A, n);
    w.labels = labels_into ? labels_into : take<int32_t>(ctx, &km_ctx::labels, A, n);
    w.L.cand_rows = take<int32_t>(ctx, &km_ctx::lists_rows, A, n);
    w.L.cand_count = take<int32_t>(ctx, &km_ctx::lists_cnt, A, n);
    w.L.cand_ids = take<int32_t>(ctx, &km_ctx::lists_ids, A, n * int64_t(MAXC));
    w.L.full_rows = take<int32_t>(ctx, &km_ctx::lists_full, A, n);
    w.L.counters = take<int32_t>(ctx, &km_ctx::lists_counters, A, 3);
    w.cn = take<F>(ctx, &km_ctx::cn, A, k);
    w.cpad = take<F>(ctx, &km_ctx::cpad, A, int64_t(k) * w.dpad);
    w.cmax_bits = take<unsigned long long>(ctx, &km_ctx::cmax, A, 4);
    // float4-readable view of X for the matrix-core kernels
    w.Xv = Xd;
    w.ldx = d;
    if (vec_view && ((d % 4) != 0 || (reinterpret_cast<uintptr_t>(Xd) % 16) != 0)) {
        const int64_t ldx = (d + 3) / 4 * 4;
        F* xp = take<F>(ctx, &km_ctx::x_pad, A, n * ldx);
        const int g = grid_for(ctx, n * ldx, 256, 8);
        hipLaunchKernelGGL(pad_rows_kernel<F>, dim3(g), dim3(256), 0, ctx->stream, Xd, n, d, ldx, xp);
        KM_LAUNCH_CHECK();
        w.Xv = xp;
        w.ldx = ldx;
    }
    // every label stays a valid cluster id from here on
    fill<int32_t>(ctx, w.labels, n, 0);
    const int grid = grid_for(ctx, n * 64, 256, 8);
    if constexpr (std::is_same<F, float>::value) {
        if (path == 3) {
            w.Xb = take<uint16_t>(ctx, &km_ctx::x_bf16, A, n * int64_t(w.dpad));
            w.enorm = take<float>(ctx, &km_ctx::enorm, A, n);
            w.xmax_bits = take<unsigned long long>(ctx, &km_ctx::xmax, A, 2);
            w.Cb = take<uint16_t>(ctx, &km_ctx::c_bf16, A, int64_t(bf16_padded_k(k)) * w.dpad);
            w.cnb = take<float>(ctx, &km_ctx::cnb, A, bf16_padded_k(k));
            w.params = take<bf16_params>(ctx, &km_ctx::params, A, 1);
            w.L.coll_rows = take<int32_t>(ctx, &km_ctx::coll_rows, A, n);
            w.L.coll_T = take<double>(ctx, &km_ctx::coll_T, A, n);
            w.L.coll_count = take<int32_t>(ctx, &km_ctx::coll_count, A, n);
            w.L.coll_ids = take<int32_t>(ctx, &km_ctx::coll_ids, A, n * int64_t(MAXC2));
            fill<unsigned long long>(ctx, w.xmax_bits, 2, 0ull);
            hipLaunchKernelGGL(convert_rows_bf16_kernel, dim3(grid), dim3(256), 0, ctx->stream, Xd, n,
                               d, w.dpad, w.Xb, w.xnorm, w.enorm, w.xmax_bits);
            KM_LAUNCH_CHECK();
            return;
        }
    }
    hipLaunchKernelGGL(row_norm_kernel<F>, dim3(grid), dim3(256), 0, ctx->stream, Xd, n, d, w.xnorm);
    KM_LAUNCH_CHECK();
}

static void collect_refine_stats(km_ctx* ctx, const int32_t* host_counters) {
    ctx->stats.refined_candidate_rows += host_counters[0];
    ctx->stats.refined_full_rows += host_counters[1];
    ctx->stats.refined_collect_rows += host_counters[2];
}

template <typename F>
static void train(km_ctx* ctx, const F* X, int64_t n, int64_t d64_, const F* C0, int64_t k64,
                  int64_t max_iter, double thr, uint32_t flags, F* C_out, int32_t* labels_out,
                  double* obj_out, int64_t* iters_out) {
    const int d = int(d64_), k = int(k64);
    const bool dev = flags & KM_FLAG_DEVICE_POINTERS;
    ctx->stats = km_stats{};
    const F* Xd = X;
    if (!dev) {
        F* xb = ctx->x_dev.get<F>(n * d);
        KM_HIP(hipMemcpyAsync(xb, X, sizeof(F) * n * d, hipMemcpyHostToDevice, ctx->stream));
        Xd = xb;
    }
    work<F> w;
    const int path = choose_path<F>(ctx, d, k, flags);
    setup_work<F>(ctx, w, Xd, n, d, k, path);
    ctx->stats.assign_path = path;
    F* Ccur = ctx->c_cur.get<F>(int64_t(k) * d);
    F* Cnext = ctx->c_next.get<F>(int64_t(k) * d);
    KM_HIP(hipMemcpyAsync(Ccur, C0, sizeof(F) * k * d,
                          dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
    int32_t* counts = ctx->counts.get<int32_t>(k);
    int64_t* offsets = ctx->offsets.get<int64_t>(k + 1);
    int32_t* chunk_base = ctx->chunk_base.get<int32_t>(k + 1);
    int32_t* empty_list = ctx->empty_list.get<int32_t>(k);
    int32_t* meta = ctx->meta.get<int32_t>(2);
    int32_t* perm = ctx->perm.get<int32_t>(n);
    double* part = ctx->part.get<double>((n / REDUCE_CHUNK + k + 1) * int64_t(d));
    int32_t* pin = ctx->pinned;
    const bool need_obj = thr > 0;
    // re-scored rows (candidate, full, collect) and refill events, summed on the device
    unsigned long long* acc = ctx->stat_acc.get<unsigned long long>(4);
    fill<unsigned long long>(ctx, acc, 4, 0ull);
    double* d64 = ctx->d64.get<double>(n);  // per-row d64 for empty clusters
    const label_sort lsort = label_sort_setup(ctx, n, k);

    hipEvent_t loop0 = nullptr, loop1 = nullptr;
    ctx->ev_used = 0;
    if (ctx->timing) {
        loop0 = ctx->event();
        loop1 = ctx->event();
        KM_HIP(hipEventRecord(loop0, ctx->stream));
    }
    double prev = double(std::numeric_limits<F>::max());
    int64_t iter = 0;
    for (iter = 0; iter < max_iter; ++iter) {
        centroid_prep<F>(ctx, w, Ccur, path);
        assign<F>(ctx, w, Ccur, path);
        label_counts(ctx, lsort, w.labels, counts);
        hipLaunchKernelGGL(scan_clusters_kernel, dim3(1), dim3(1024), 0, ctx->stream, counts, k,
                           offsets, chunk_base, empty_list, meta);
        KM_LAUNCH_CHECK();
        label_group(ctx, lsort, w.labels, offsets, perm);
        launch_reduce<F>(ctx, w, perm, offsets, chunk_base, meta, part);
        {
            const int grid = grid_for(ctx, int64_t(k) * d, 256, 4);
            hipLaunchKernelGGL(finalize_centroids_kernel<F>, dim3(grid), dim3(256), 0, ctx->stream,
                               part, counts, chunk_base, k, d, Cnext);
            KM_LAUNCH_CHECK();
        }
        double* objp = nullptr;
        if (need_obj) objp = objective<F>(ctx, w, Ccur, d64);
        // empty clusters and the statistics stay on the device: with no stop
        // threshold an iteration needs no host round trip
        empty_clusters<F>(ctx, w, Ccur, Cnext, empty_list, meta, objp, d64, acc);
        std::swap(Ccur, Cnext);
        if (need_obj) {
            double obj = 0.0;
            KM_HIP(hipMemcpyAsync(pin + 6, objp, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            KM_HIP(hipStreamSynchronize(ctx->stream));
            std::memcpy(&obj, pin + 6, sizeof(double));
            if (obj + thr > prev) {
                iter++;
                break;
            }
            prev = obj;
        }
    }
    // final assignment with the final centroids and the objective
    centroid_prep<F>(ctx, w, Ccur, path);
    assign<F>(ctx, w, Ccur, path);
    double* objp = objective<F>(ctx, w, Ccur, nullptr);
    if (ctx->timing) KM_HIP(hipEventRecord(loop1, ctx->stream));
    KM_HIP(hipMemcpyAsync(pin + 2, w.L.counters, 3 * sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    KM_HIP(hipMemcpyAsync(pin + 6, objp, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    unsigned long long acc_h[4];
    KM_HIP(hipMemcpyAsync(acc_h, acc, sizeof(acc_h), hipMemcpyDeviceToHost, ctx->stream));
    const hipMemcpyKind out_kind = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    KM_HIP(hipMemcpyAsync(C_out, Ccur, sizeof(F) * k * d, out_kind, ctx->stream));
    KM_HIP(hipMemcpyAsync(labels_out, w.labels, sizeof(int32_t) * n, out_kind, ctx->stream));
    KM_HIP(hipStreamSynchronize(ctx->stream));
    collect_refine_stats(ctx, pin + 2);
    ctx->stats.refined_candidate_rows += int64_t(acc_h[0]);
    ctx->stats.refined_full_rows += int64_t(acc_h[1]);
    ctx->stats.refined_collect_rows += int64_t(acc_h[2]);
    ctx->stats.empty_cluster_events += int64_t(acc_h[3]);
    double obj;
    std::memcpy(&obj, pin + 6, sizeof(double));
    if (obj_out) *obj_out = obj;
    if (iters_out) *iters_out = iter;
    if (ctx->timing) {
        float ms = 0.f;
        KM_HIP(hipEventElapsedTime(&ms, loop0, loop1));
        ctx->stats.iterations_ms = ms;
        double tot = 0.0;
        for (size_t i = 2; i + 1 < ctx->ev_used; i += 2) {
            float a = 0.f;
            KM_HIP(hipEventElapsedTime(&a, ctx->ev_pool[i], ctx->ev_pool[i + 1]));
            tot += a;
        }
        ctx->stats.assign_kernel_ms = tot;
    }
}

template <typename F>
static void infer(km_ctx* ctx, const F* X, int64_t n, int64_t d64_, const F* C, int64_t k64,
                  uint32_t flags, int32_t* labels_out, double* obj_out) {
    const int d = int(d64_), k = int(k64);
    const bool dev = flags & KM_FLAG_DEVICE_POINTERS;
    ctx->stats = km_stats{};
    ctx->ev_used = 0;
    const F* Xd = X;
    if (!dev) {
        F* xb = ctx->x_dev.get<F>(n * d);
        KM_HIP(hipMemcpyAsync(xb, X, sizeof(F) * n * d, hipMemcpyHostToDevice, ctx->stream));
        Xd = xb;
    }
    work<F> w;
    const int path = choose_path<F>(ctx, d, k, flags);
    setup_work<F>(ctx, w, Xd, n, d, k, path);
    ctx->stats.assign_path = path;
    F* Cd = ctx->c_cur.get<F>(int64_t(k) * d);
    KM_HIP(hipMemcpyAsync(Cd, C, sizeof(F) * k * d,
                          dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
    centroid_prep<F>(ctx, w, Cd, path);
    assign<F>(ctx, w, Cd, path);
    int32_t* pin = ctx->pinned;
    const bool want_obj = flags & KM_FLAG_EXACT_OBJECTIVE;
    if (want_obj) {
        double* objp = objective<F>(ctx, w, Cd, nullptr);
        KM_HIP(hipMemcpyAsync(pin + 6, objp, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    }
    KM_HIP(hipMemcpyAsync(pin + 2, w.L.counters, 3 * sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    KM_HIP(hipMemcpyAsync(labels_out, w.labels, sizeof(int32_t) * n,
                          dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->stream));
    KM_HIP(hipStreamSynchronize(ctx->stream));
    collect_refine_stats(ctx, pin + 2);
    if (want_obj && obj_out) std::memcpy(obj_out, pin + 6, sizeof(double));
    if (ctx->timing) {
        double tot = 0.0;
        for (size_t i = 0; i + 1 < ctx->ev_used; i += 2) {
            float a = 0.f;
            KM_HIP(hipEventElapsedTime(&a, ctx->ev_pool[i], ctx->ev_pool[i + 1]));
            tot += a;
        }
        ctx->stats.assign_kernel_ms = tot;
    }
}

// ---------------------------------------------------------------------------
// Inference without allocation (km_lloyd_infer_f32_noalloc).  Every workspace
// piece comes from the device's fixed scratch array; rows go through in blocks
// that fit it, each block a whole number of objective chunks (OBJ_CHUNK rows),
// so the chunk sums and their in-order addition are those of infer().  The
// labels are the exact-form argmin on every path, so they do not depend on the
// blocking either.  Calls that overlap on one device share the array and race.
// ---------------------------------------------------------------------------
constexpr int64_t NOALLOC_SLACK = 16384;  // alignment of the arena's pieces

// arena bytes of one block of R rows (mirrors setup_work, launch_bf16 and the
// objective chunks); the arena itself refuses anything beyond its end
template <typename F>
static size_t noalloc_block_bytes(const km_ctx* ctx, int path, int64_t R, int d, int k, bool pad_x,
                                  bool want_obj) {
    const int64_t dpad = pad_dim(d);
    int64_t b = R * (4 + 4 + 4 + 4 * int64_t(MAXC) + 4) + 16 + int64_t(sizeof(F)) * k +
                int64_t(sizeof(F)) * k * dpad + 32;
    if (path == 2 && pad_x) b += R * ((d + 3) / 4 * 4) * int64_t(sizeof(F));
    if (path == 3) {
        const int64_t kp = bf16_padded_k(k);
        b += R * (2 * dpad + 4 + 4 + 8 + 4 + 4 * int64_t(MAXC2)) + 16 + 2 * kp * dpad + 4 * kp + 64;
        const int64_t rows_per_wg = dpad >= 256 ? 128 : 256;  // launch_bf16: 4 waves x 32 x NB
        const int64_t tiles = (R + rows_per_wg - 1) / rows_per_wg;
        const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(tiles, int64_t(ctx->cu_count) * 2));
        const int64_t ents = (tiles + grid - 1) / grid * rows_per_wg * grid;
        b += ents * (4 + 4 + 4 * int64_t(MAXC) + 4 + 8) + 2 * 4 * grid * 4;  // + segment counts
    }
    if (want_obj) b += 8 * ((R + OBJ_CHUNK - 1) / OBJ_CHUNK);
    return size_t(b + NOALLOC_SLACK);
}

static void infer_noalloc(km_ctx* ctx, const float* X, int64_t n, int64_t d64_, const float* C,
                          int64_t k64, uint32_t flags, int32_t* labels_out, double* obj_out) {
    using F = float;
    const int d = int(d64_), k = int(k64);
    if (!(flags & KM_FLAG_DEVICE_POINTERS))
        throw status_error(KM_INVALID_ARGUMENT,
                           "km_lloyd_infer_f32_noalloc takes device pointers only (KM_FLAG_DEVICE_POINTERS)");
    if (!ctx->scratch)
        throw status_error(KM_INTERNAL_ERROR,
                           "the fixed scratch buffer is not available on this device: " + ctx->scratch_err);
    ctx->stats = km_stats{};
    const int path = choose_path<F>(ctx, d, k, flags);
    const bool want_obj = flags & KM_FLAG_EXACT_OBJECTIVE;
    const bool pad_x = (d % 4) != 0 || (reinterpret_cast<uintptr_t>(X) % 16) != 0;
    const size_t cap = NOALLOC_BYTES - 256;  // the objective accumulator sits in front
    // largest block: a multiple of OBJ_CHUNK rows (or all rows) that fits
    int64_t R = std::min<int64_t>(n, int64_t(1) << 22);
    R = R < n ? R / OBJ_CHUNK * OBJ_CHUNK : R;
    while (R > OBJ_CHUNK && noalloc_block_bytes<F>(ctx, path, R, d, k, pad_x, want_obj) > cap)
        R = (R - 1) / OBJ_CHUNK * OBJ_CHUNK;
    if (noalloc_block_bytes<F>(ctx, path, R, d, k, pad_x, want_obj) > cap)
        throw status_error(KM_INVALID_ARGUMENT,
                           "k x d too large for km_lloyd_infer_f32_noalloc: one block of " +
                               std::to_string(R) + " rows needs " +
                               std::to_string(noalloc_block_bytes<F>(ctx, path, R, d, k, pad_x, want_obj)) +
                               " bytes of the " + std::to_string(NOALLOC_BYTES) +
                               "-byte scratch buffer");
    // no timing events here: creating one may allocate
    struct timing_off {
        km_ctx* c;
        bool was;
        ~timing_off() { c->timing = was; }
    } restore{ctx, ctx->timing};
    ctx->timing = false;
    ctx->stats.assign_path = path;
    unsigned char* base = static_cast<unsigned char*>(ctx->scratch);
    double* acc = reinterpret_cast<double*>(base);
    if (want_obj) fill<double>(ctx, acc, 1, 0.0);
    int32_t* pin = ctx->pinned;
    for (int64_t r0 = 0; r0 < n; r0 += R) {
        const int64_t rows = std::min<int64_t>(R, n - r0);
        arena A{base + 256, cap, 0};
        work<F> w;
        setup_work<F>(ctx, w, X + r0 * d, rows, d, k, path, &A, labels_out + r0);
        centroid_prep<F>(ctx, w, C, path);
        assign<F>(ctx, w, C, path);
        if (want_obj) {
            const int64_t nchunks = (rows + OBJ_CHUNK - 1) / OBJ_CHUNK;
            double* chunks = A.get<double>(nchunks);
            const int grid = int(std::min<int64_t>(nchunks, int64_t(ctx->cu_count) * 8));
            launch_objective<F>(ctx, grid, w.X, w.n, w.d, C, w.labels, nullptr, chunks);
            hipLaunchKernelGGL(sum_sequential_into_kernel, dim3(1), dim3(64), 0, ctx->stream, chunks,
                               nchunks, acc);
            KM_LAUNCH_CHECK();
        }
        // the next block reuses the arena: its pieces are stream-ordered after this
        // block's kernels, and the counters are read before they are cleared again
        KM_HIP(hipMemcpyAsync(pin + 2, w.L.counters, 3 * sizeof(int32_t), hipMemcpyDeviceToHost,
                              ctx->stream));
        KM_HIP(hipStreamSynchronize(ctx->stream));
        collect_refine_stats(ctx, pin + 2);
    }
    if (want_obj) {
        KM_HIP(hipMemcpyAsync(pin + 6, acc, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KM_HIP(hipStreamSynchronize(ctx->stream));
        if (obj_out) std::memcpy(obj_out, pin + 6, sizeof(double));
    }
}

// ---------------------------------------------------------------------------
// CSR Lloyd (method lloyd_csr: train_kernel_lloyd_csr_dpc.cpp:76-242, and the
// lloyd_csr branch of infer_kernel_dpc.cpp).  Loop, stop rule, grouping by label,
// chunk sums, finalize and empty-cluster policy are the dense driver's; the
// distances are the expanded form over the stored entries (km_csr.hpp).
// ---------------------------------------------------------------------------
template <typename F>
struct csr_view {
    const F* vals;
    const int64_t* cols;
    const int64_t* offs;
    int64_t base, n, d, nnz;
};

// device view of the caller's CSR arrays (uploaded unless they are device
// pointers), after the structure check; *increasing (if given) tells whether
// every row's column indices increase strictly
template <typename F>
static csr_view<F> csr_prepare(km_ctx* ctx, const F* vals, const int64_t* cols,
                               const int64_t* offs, int64_t n, int64_t d, int64_t nnz,
                               int32_t indexing, bool dev, bool* increasing = nullptr) {
    csr_view<F> v{vals, cols, offs, indexing == KM_SPARSE_ONE_BASED ? 1 : 0, n, d, nnz};
    if (!dev) {
        F* vb = ctx->csr_vals.get<F>(std::max<int64_t>(nnz, 1));
        int64_t* cb = ctx->csr_cols.get<int64_t>(std::max<int64_t>(nnz, 1));
        int64_t* ob = ctx->csr_offs.get<int64_t>(n + 1);
        if (nnz > 0) {
            KM_HIP(hipMemcpyAsync(vb, vals, sizeof(F) * nnz, hipMemcpyHostToDevice, ctx->stream));
            KM_HIP(hipMemcpyAsync(cb, cols, sizeof(int64_t) * nnz, hipMemcpyHostToDevice,
                                  ctx->stream));
        }
        KM_HIP(hipMemcpyAsync(ob, offs, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice,
                              ctx->stream));
        v.vals = vb;
        v.cols = cb;
        v.offs = ob;
    } else if (nnz == 0) {
        v.vals = ctx->csr_vals.get<F>(1);
        v.cols = ctx->csr_cols.get<int64_t>(1);
    }
    int32_t* flag = ctx->csr_flag.get<int32_t>(1);
    fill<int32_t>(ctx, flag, 1, 0);
    hipLaunchKernelGGL(csr_check_kernel, dim3(grid_for(ctx, n, 256, 4)), dim3(256), 0, ctx->stream,
                       v.cols, v.offs, v.base, n, d, nnz, flag);
    KM_LAUNCH_CHECK();
    int32_t* pin = ctx->pinned;
    KM_HIP(hipMemcpyAsync(pin, flag, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    KM_HIP(hipStreamSynchronize(ctx->stream));
    if (pin[0] & 1)
        throw status_error(KM_INVALID_ARGUMENT,
                           "Row offsets of the sparse data must be non-decreasing, from the "
                           "indexing base to base + non-zero count");
    if (pin[0] & 2)
        throw status_error(KM_INVALID_ARGUMENT,
                           "Column indices of the sparse data must lie in [base, base + column count)");
    if (increasing) *increasing = (pin[0] & 4) == 0;
    return v;
}

// labels and closest distances of every row against C (k x d)
template <typename F>
static void csr_assign(km_ctx* ctx, const csr_view<F>& x, const F* C, int k, const F* xs,
                       int32_t* labels, F* closest) {
    constexpr int VEC = csr_vec<F>::N;
    const int kp = (k + VEC - 1) / VEC * VEC;  // Ct row stride: whole 16-byte vectors
    F* ct = ctx->csr_ct.get<F>(int64_t(kp) * x.d);
    F* cs = ctx->csr_cs.get<F>(k);
    hipLaunchKernelGGL(csr_transpose_kernel<F>,
                       dim3(unsigned((x.d + 63) / 64), unsigned(std::min((kp + 63) / 64, 65535))),
                       dim3(256), 0, ctx->stream, C, int64_t(k), int64_t(kp), x.d, ct);
    KM_LAUNCH_CHECK();
    hipLaunchKernelGGL(csr_centroid_squares_kernel<F>, dim3(unsigned(std::min(k, ctx->cu_count * 8))),
                       dim3(64), size_t(x.d) * sizeof(F), ctx->stream, C, int64_t(k), int(x.d), cs);
    KM_LAUNCH_CHECK();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->timing) {
        e0 = ctx->event();
        e1 = ctx->event();
        KM_HIP(hipEventRecord(e0, ctx->stream));
    }
    const int grid = grid_for(ctx, x.n * 64, 256, 8);
#define KM_CSRA(JT_)                                                                                \
    hipLaunchKernelGGL((csr_assign_kernel<F, JT_>), dim3(grid), dim3(256), 0, ctx->stream, x.vals,   \
                       x.cols, x.offs, x.base, x.n, k, kp, ct, cs, xs, labels, closest)
    // centroids per lane and block: VEC * JT
    if (k <= 64 * VEC) KM_CSRA(1);
    else if (k <= 128 * VEC) KM_CSRA(2);
    else KM_CSRA(4);
#undef KM_CSRA
    KM_LAUNCH_CHECK();
    if (ctx->timing) KM_HIP(hipEventRecord(e1, ctx->stream));
    ctx->stats.assign_launches++;
}

// objective = sum of closest (as double) in the fixed order; d64 per row if asked
template <typename F>
static double* csr_objective(km_ctx* ctx, const F* closest, int64_t n, double* d64_out,
                             const int32_t* gate = nullptr) {
    const int64_t nchunks = (n + OBJ_CHUNK - 1) / OBJ_CHUNK;
    double* chunks = ctx->obj_chunks.get<double>(nchunks);
    double* out = ctx->obj_dev.get<double>(1);
    const int grid = int(std::min<int64_t>(nchunks, int64_t(ctx->cu_count) * 8));
    hipLaunchKernelGGL(csr_objective_chunks_kernel<F>, dim3(grid), dim3(256), 0, ctx->stream,
                       closest, n, d64_out, chunks, gate);
    KM_LAUNCH_CHECK();
    if (gate) return nullptr;  // the d64 pass of the empty-cluster refill only
    hipLaunchKernelGGL(sum_sequential_kernel, dim3(1), dim3(64), 0, ctx->stream, chunks, nchunks, out);
    KM_LAUNCH_CHECK();
    return out;
}

static void csr_timing_stats(km_ctx* ctx, size_t first_pair) {
    double tot = 0.0;
    for (size_t i = first_pair; i + 1 < ctx->ev_used; i += 2) {
        float a = 0.f;
        KM_HIP(hipEventElapsedTime(&a, ctx->ev_pool[i], ctx->ev_pool[i + 1]));
        tot += a;
    }
    ctx->stats.assign_kernel_ms = tot;
}

template <typename F>
static void train_csr(km_ctx* ctx, const F* vals, const int64_t* cols, const int64_t* offs,
                      int64_t n, int64_t d64_, int64_t nnz, int32_t indexing, const F* C0,
                      int64_t k64, int64_t max_iter, double thr, uint32_t flags, F* C_out,
                      int32_t* labels_out, double* obj_out, int64_t* iters_out) {
    const int d = int(d64_), k = int(k64);
    const bool dev = flags & KM_FLAG_DEVICE_POINTERS;
    ctx->stats = km_stats{};
    ctx->ev_used = 0;
    hipEvent_t loop0 = nullptr, loop1 = nullptr;
    if (ctx->timing) {
        loop0 = ctx->event();
        loop1 = ctx->event();
    }
    const csr_view<F> x = csr_prepare<F>(ctx, vals, cols, offs, n, d, nnz, indexing, dev);
    F* xs = ctx->csr_xs.get<F>(n);
    F* closest = ctx->csr_closest.get<F>(n);
    int32_t* labels = ctx->labels.get<int32_t>(n);
    hipLaunchKernelGGL(csr_row_squares_kernel<F>, dim3(grid_for(ctx, n, 256, 4)), dim3(256), 0,
                       ctx->stream, x.vals, x.offs, x.base, n, xs);
    KM_LAUNCH_CHECK();
    F* Ccur = ctx->c_cur.get<F>(int64_t(k) * d);
    F* Cnext = ctx->c_next.get<F>(int64_t(k) * d);
    KM_HIP(hipMemcpyAsync(Ccur, C0, sizeof(F) * k * d,
                          dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
    int32_t* counts = ctx->counts.get<int32_t>(k);
    int64_t* offsets = ctx->offsets.get<int64_t>(k + 1);
    int32_t* chunk_base = ctx->chunk_base.get<int32_t>(k + 1);
    int32_t* empty_list = ctx->empty_list.get<int32_t>(k);
    int32_t* meta = ctx->meta.get<int32_t>(2);
    int32_t* perm = ctx->perm.get<int32_t>(n);
    double* part = ctx->part.get<double>((n / REDUCE_CHUNK + k + 1) * int64_t(d));
    int32_t* pin = ctx->pinned;
    const bool need_obj = thr > 0;
    const size_t lds = size_t(reduce_groups(d)) * d * sizeof(double);
    KM_HIP(hipFuncSetAttribute((const void*)csr_chunk_sums_kernel<F>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    double* d64 = ctx->d64.get<double>(n);  // refill candidates' keys
    const label_sort lsort = label_sort_setup(ctx, n, k);
    unsigned long long* acc = ctx->stat_acc.get<unsigned long long>(4);  // [3]: refill events
    fill<unsigned long long>(ctx, acc, 4, 0ull);
    if (ctx->timing) KM_HIP(hipEventRecord(loop0, ctx->stream));

    double prev = double(std::numeric_limits<F>::max());
    int64_t iter = 0;
    for (iter = 0; iter < max_iter; ++iter) {
        csr_assign<F>(ctx, x, Ccur, k, xs, labels, closest);
        label_counts(ctx, lsort, labels, counts);
        hipLaunchKernelGGL(scan_clusters_kernel, dim3(1), dim3(1024), 0, ctx->stream, counts, k,
                           offsets, chunk_base, empty_list, meta);
        KM_LAUNCH_CHECK();
        label_group(ctx, lsort, labels, offsets, perm);
        const int64_t max_chunks = n / REDUCE_CHUNK + k + 1;
        hipLaunchKernelGGL(csr_chunk_sums_kernel<F>, dim3(grid_for(ctx, max_chunks * 64, 64, 16)),
                           dim3(64), lds, ctx->stream, x.vals, x.cols, x.offs, x.base, d, k, perm,
                           offsets, chunk_base, meta, part);
        KM_LAUNCH_CHECK();
        hipLaunchKernelGGL(finalize_centroids_kernel<F>, dim3(grid_for(ctx, int64_t(k) * d, 256, 4)),
                           dim3(256), 0, ctx->stream, part, counts, chunk_base, k, d, Cnext);
        KM_LAUNCH_CHECK();
        // empty clusters on the device (d64 pass gated on the empty count, the
        // selection kernels, the sparse fill): no host round trip without a
        // stop threshold
        double* objp = nullptr;
        if (need_obj)
            objp = csr_objective<F>(ctx, closest, n, d64);
        else
            csr_objective<F>(ctx, closest, n, d64, meta);
        int32_t* cand = launch_empty_clusters<F>(ctx, (const F*)nullptr, n, d, k, meta, empty_list,
                                                 (const int32_t*)nullptr, Cnext, d64, objp, acc);
        hipLaunchKernelGGL(csr_fill_empty_kernel<F>, dim3(std::min(k, ctx->cu_count * 4)), dim3(64), 0,
                           ctx->stream, x.vals, x.cols, x.offs, x.base, d, empty_list, cand, meta, Cnext);
        KM_LAUNCH_CHECK();
        std::swap(Ccur, Cnext);
        if (need_obj) {
            double obj = 0.0;
            KM_HIP(hipMemcpyAsync(pin + 6, objp, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            KM_HIP(hipStreamSynchronize(ctx->stream));
            std::memcpy(&obj, pin + 6, sizeof(double));
            if (obj + thr > prev) {
                iter++;
                break;
            }
            prev = obj;
        }
    }
    csr_assign<F>(ctx, x, Ccur, k, xs, labels, closest);
    double* objp = csr_objective<F>(ctx, closest, n, nullptr);
    if (ctx->timing) KM_HIP(hipEventRecord(loop1, ctx->stream));
    KM_HIP(hipMemcpyAsync(pin + 6, objp, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    unsigned long long acc_h[4];
    KM_HIP(hipMemcpyAsync(acc_h, acc, sizeof(acc_h), hipMemcpyDeviceToHost, ctx->stream));
    const hipMemcpyKind out_kind = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    KM_HIP(hipMemcpyAsync(C_out, Ccur, sizeof(F) * k * d, out_kind, ctx->stream));
    KM_HIP(hipMemcpyAsync(labels_out, labels, sizeof(int32_t) * n, out_kind, ctx->stream));
    KM_HIP(hipStreamSynchronize(ctx->stream));
    ctx->stats.empty_cluster_events += int64_t(acc_h[3]);
    double obj;
    std::memcpy(&obj, pin + 6, sizeof(double));
    if (obj_out) *obj_out = obj;
    if (iters_out) *iters_out = iter;
    if (ctx->timing) {
        float ms = 0.f;
        KM_HIP(hipEventElapsedTime(&ms, loop0, loop1));
        ctx->stats.iterations_ms = ms;
        csr_timing_stats(ctx, 2);
    }
}

template <typename F>
static void infer_csr(km_ctx* ctx, const F* vals, const int64_t* cols, const int64_t* offs,
                      int64_t n, int64_t d64_, int64_t nnz, int32_t indexing, const F* C, int64_t k64,
                      uint32_t flags, int32_t* labels_out, double* obj_out) {
    const int d = int(d64_), k = int(k64);
    const bool dev = flags & KM_FLAG_DEVICE_POINTERS;
    ctx->stats = km_stats{};
    ctx->ev_used = 0;
    const csr_view<F> x = csr_prepare<F>(ctx, vals, cols, offs, n, d, nnz, indexing, dev);
    F* xs = ctx->csr_xs.get<F>(n);
    F* closest = ctx->csr_closest.get<F>(n);
    int32_t* labels = ctx->labels.get<int32_t>(n);
    hipLaunchKernelGGL(csr_row_squares_kernel<F>, dim3(grid_for(ctx, n, 256, 4)), dim3(256), 0,
                       ctx->stream, x.vals, x.offs, x.base, n, xs);
    KM_LAUNCH_CHECK();
    F* Cd = ctx->c_cur.get<F>(int64_t(k) * d);
    KM_HIP(hipMemcpyAsync(Cd, C, sizeof(F) * k * d,
                          dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
    csr_assign<F>(ctx, x, Cd, k, xs, labels, closest);
    int32_t* pin = ctx->pinned;
    const bool want_obj = flags & KM_FLAG_EXACT_OBJECTIVE;
    if (want_obj) {
        double* objp = csr_objective<F>(ctx, closest, n, nullptr);
        KM_HIP(hipMemcpyAsync(pin + 6, objp, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    }
    KM_HIP(hipMemcpyAsync(labels_out, labels, sizeof(int32_t) * n,
                          dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->stream));
    KM_HIP(hipStreamSynchronize(ctx->stream));
    if (want_obj && obj_out) std::memcpy(obj_out, pin + 6, sizeof(double));
    if (ctx->timing) csr_timing_stats(ctx, 0);
}

// --- precondition checks (reference order: train_ops.hpp:40-61, common.cpp:82-115) ---
// ---------------------------------------------------------------------------
// Batched plans (km_batch_train_f32, km_batch_step_f32).  A plan is a set of
// instances with its instance table and prefix tables, built on the host; each
// stage of an iteration is one launch over every instance of the plan, so an
// iteration needs no host synchronisation and can be captured as a HIP graph.
// ---------------------------------------------------------------------------
struct batch_plan {
    // prefix tables: assignment batches, histogram chunks, centroid elements,
    // d64 batches, objective chunks
    enum { P_CNT = 0, P_EL, P_D64, P_OBJ, P_NUM };
    // one assignment launch per tile variant in use: its instances
    // (asel[sel_off ..]) and the prefix of their workgroup counts (apre[pre_off ..])
    struct asg_launch {
        int v;
        int nsel;
        int64_t sel_off, pre_off, batches;
    };
    int B = 0;
    bool train = false;
    int64_t rows = 0, cfl = 0, kk = 0, pp = 0, oo = 0, hh = 0;
    unsigned kbits = 1, lbits = 0;
    int maxk = 0;
    // rows grouped by label with the counting sort (every k <= BATCH_HIST_MAX), or
    // with one radix sort of (instance, label) keys
    bool csort = false;
    // instances summed by batch_reduce_vec_kernel / batch_reduce_kernel
    bool any_vec = false, any_scalar = false;
    std::vector<batch_inst> hi;
    std::vector<int64_t> pre;
    std::vector<int32_t> asel;
    std::vector<int64_t> apre;
    std::vector<asg_launch> launches;
    // device views (plan_upload)
    batch_inst* dI = nullptr;
    int64_t* dpre = nullptr;
    int32_t* d_asel = nullptr;
    int64_t* d_apre = nullptr;
    int32_t* labels = nullptr;
    uint32_t* skeys = nullptr;
    int32_t* perm = nullptr;
    double* d64 = nullptr;
    float* Ccur = nullptr;
    float* Cnext = nullptr;
    int32_t* counts = nullptr;
    int64_t* offsets = nullptr;
    int32_t* cb = nullptr;
    int32_t* empty = nullptr;
    int32_t* meta = nullptr;
    int32_t* bhist = nullptr;
    uint32_t* ticket = nullptr;
    int64_t* qpre = nullptr;
    int64_t* d64pre = nullptr;  // d64 blocks of the instances with empty clusters (device prefix)
    double* part = nullptr;
    double* objc = nullptr;
    double* objd = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;

    int64_t P(int t, int i) const { return pre[size_t(t) * (B + 1) + i]; }
    const int64_t* DP(int t) const { return dpre + size_t(t) * (B + 1); }
};

// Instance table and prefix tables of `count` instances.  train: the instances
// are iterated (training preconditions, C_out required); otherwise they only get
// a final pass (inference preconditions; labels_out required, C_out optional).
// (d, k) bucket of the tile table: d in {<=8, <=16, <=32, <=64, <=128, <=256, more},
// k in {<=8, <=16, <=32, <=64, <=128, more}; index 6 * d_bucket + k_bucket
static int tile_bucket(int64_t d, int64_t k) {
    const int db = d <= 8 ? 0 : d <= 16 ? 1 : d <= 32 ? 2 : d <= 64 ? 3 : d <= 128 ? 4 : d <= 256 ? 5 : 6;
    const int kb = k <= 8 ? 0 : k <= 16 ? 1 : k <= 32 ? 2 : k <= 64 ? 3 : k <= 128 ? 4 : 5;
    return db * 6 + kb;
}

// untuned choice: the narrowest centroid tile that still holds k
static int default_tile(int64_t k) { return k <= 8 ? 4 : k <= 16 ? 3 : k <= 32 ? 2 : 0; }

static bool mfma_tile_ok(int64_t d, int64_t k) { return d <= 128 && k <= BATCH_MFMA_MAXK; }

// tile variant of an instance (the matrix-core variant only where it applies)
static int tile_for(const km_ctx* ctx, int64_t d, int64_t k) {
    int t = ctx->tile_force > 0 ? ctx->tile_force - 1 : ctx->tile_table[tile_bucket(d, k)];
    if (t < 0 || (t == BATCH_TILE_MFMA && !mfma_tile_ok(d, k))) t = default_tile(k);
    return t;
}

// launch kind of a tile variant: the matrix-core variant splits by padded width
static int launch_kind(int v, int64_t d) {
    if (v != BATCH_TILE_MFMA) return v;
    return d <= 16 ? BATCH_TILE_MFMA + 3 : BATCH_TILE_MFMA + (d <= 32 ? 0 : d <= 64 ? 1 : 2);
}

static void plan_build(batch_plan& p, const km_ctx* ctx, const km_batch_instance_f32* inst, int64_t count,
                       bool train) {
    if (!inst || count <= 0) throw status_error(KM_DOMAIN_ERROR, "Input batch is empty");
    if (count > std::numeric_limits<int32_t>::max())
        throw status_error(KM_DOMAIN_ERROR, "Instance count is greater than max int32 value");
    const int B = int(count);
    p.B = B;
    p.train = train;
    p.hi.assign(B, batch_inst{});
    p.pre.assign(size_t(batch_plan::P_NUM) * (B + 1), 0);
    auto P = [&](int t, int i) -> int64_t& { return p.pre[size_t(t) * (B + 1) + i]; };
    int64_t rows = 0, cfl = 0, kk = 0, pp = 0, oo = 0, hh = 0;
    int maxk = 1;
    for (int i = 0; i < B; ++i) {
        const km_batch_instance_f32& q = inst[i];
        if (!q.X || !q.C0 || (train && !q.C_out) || (!train && !q.labels_out))
            throw status_error(KM_INVALID_ARGUMENT, "Batch instance " + std::to_string(i) +
                                                        ": input and output buffers are required");
        if (q.n <= 0 || q.d <= 0) throw status_error(KM_DOMAIN_ERROR, "Input data is empty");
        if (q.k <= 0 || q.k > std::numeric_limits<int32_t>::max())
            throw status_error(KM_DOMAIN_ERROR, "Cluster count should be > 0 and <= max int32 value");
        if (q.n > std::numeric_limits<int32_t>::max())
            throw status_error(KM_DOMAIN_ERROR, "Row count is greater than max int32 value");
        if (train && q.k > q.n)
            throw status_error(KM_INVALID_ARGUMENT,
                               "Cluster count should be less than or equal to the row count");
        if (q.d > 4096) throw status_error(KM_UNIMPLEMENTED, "column count above 4096 is not supported");
        batch_inst& b = p.hi[i];
        b.X = q.X;
        b.C0 = q.C0;
        b.C_out = q.C_out;
        b.labels_out = q.labels_out;
        b.n = q.n;
        b.d = int32_t(q.d);
        b.k = int32_t(q.k);
        b.index = i;
        b.vec = (q.d % 4 == 0 && reinterpret_cast<uintptr_t>(q.X) % 16 == 0) ? 1 : 0;
        if (b.vec && q.d <= 256)
            p.any_vec = true;
        else
            p.any_scalar = true;
        b.row0 = rows;
        b.c0 = cfl;
        b.k0 = kk;
        b.p0 = pp;
        b.o0 = oo;
        b.h0 = hh;
        rows += q.n;
        cfl += (q.k * q.d + 3) / 4 * 4;
        kk += q.k;
        pp += (q.n / REDUCE_CHUNK + q.k + 1) * q.d;
        oo += (q.n + OBJ_CHUNK - 1) / OBJ_CHUNK;
        hh += (q.n + BATCH_CNT_ROWS - 1) / BATCH_CNT_ROWS * q.k;
        maxk = std::max<int>(maxk, int(q.k));
        P(batch_plan::P_CNT, i + 1) = P(batch_plan::P_CNT, i) + (q.n + BATCH_CNT_ROWS - 1) / BATCH_CNT_ROWS;
        P(batch_plan::P_EL, i + 1) = P(batch_plan::P_EL, i) + q.k * q.d;
        P(batch_plan::P_D64, i + 1) = P(batch_plan::P_D64, i) + (q.n + BATCH_D64_ROWS - 1) / BATCH_D64_ROWS;
        P(batch_plan::P_OBJ, i + 1) = P(batch_plan::P_OBJ, i) + (q.n + OBJ_CHUNK - 1) / OBJ_CHUNK;
    }
    if (rows >= (int64_t(1) << 31))
        throw status_error(KM_UNIMPLEMENTED, "a batch holds at most 2^31 - 1 rows in total");
    p.rows = rows;
    p.cfl = cfl;
    p.kk = kk;
    p.pp = pp;
    p.oo = oo;
    p.hh = hh;
    p.maxk = maxk;
    p.csort = train && maxk <= BATCH_HIST_MAX;
    // rows are grouped by (instance, label) with one stable radix sort of the keys
    // (index << lbits) | label; instances are contiguous, so each instance's rows
    // come out by label in row order
    unsigned lbits = 0, ibits = 0;
    while ((1ull << lbits) < (unsigned long long)maxk) lbits++;
    while ((1ull << ibits) < (unsigned long long)B) ibits++;
    if (train && lbits + ibits > 32)
        throw status_error(KM_UNIMPLEMENTED, "batch too large for 32-bit sort keys "
                                             "(instances x max cluster count above 2^32)");
    p.lbits = lbits;
    p.kbits = std::max(1u, lbits + ibits);
    // assignment launches: instances grouped by tile variant, in instance order
    std::vector<int> var(B);
    for (int i = 0; i < B; ++i) var[i] = launch_kind(tile_for(ctx, inst[i].d, inst[i].k), inst[i].d);
    p.asel.clear();
    p.apre.clear();
    p.launches.clear();
    for (int v = 0; v < BATCH_LAUNCH_KINDS; ++v) {
        batch_plan::asg_launch L{v, 0, int64_t(p.asel.size()), int64_t(p.apre.size()), 0};
        const int64_t rb = batch_kind_rows(v);
        p.apre.push_back(0);
        for (int i = 0; i < B; ++i) {
            if (var[i] != v) continue;
            p.asel.push_back(i);
            L.batches += (inst[i].n + rb - 1) / rb;
            p.apre.push_back(L.batches);
            L.nsel++;
        }
        if (L.nsel == 0) {
            p.apre.pop_back();
            continue;
        }
        p.launches.push_back(L);
    }
}

static int batch_grid(const km_ctx* ctx, int64_t items) {
    const int64_t cap = int64_t(ctx->cu_count) * 8;
    return int(std::max<int64_t>(1, std::min<int64_t>(items, cap)));
}

// Device workspace of the plan and the table upload (stream-ordered: the host
// tables must stay alive until the stream has passed the copies).
static void plan_upload(km_ctx* ctx, batch_ws& ws, batch_plan& p) {
    const int B = p.B;
    p.dI = ws.inst.get<batch_inst>(B);
    p.dpre = ws.pre.get<int64_t>(p.pre.size());
    p.Ccur = ws.ccur.get<float>(p.cfl);
    p.objc = ws.objc.get<double>(p.oo);
    p.objd = ws.obj.get<double>(B);
    p.d_asel = ws.asel.get<int32_t>(p.asel.size());
    p.d_apre = ws.apre.get<int64_t>(p.apre.size());
    p.d64 = ws.d64.get<double>(p.rows);  // the final pass's objective, and empty clusters
    if (p.train) {
        p.labels = ws.labels.get<int32_t>(p.rows);
        if (!p.csort) p.skeys = ws.keys.get<uint32_t>(2 * p.rows);  // keys, then their sorted copy
        p.perm = ws.perm.get<int32_t>(p.rows);
        p.Cnext = ws.cnext.get<float>(p.cfl);
        p.counts = ws.counts.get<int32_t>(p.kk);
        p.offsets = ws.offsets.get<int64_t>(p.kk + B);
        p.cb = ws.cb.get<int32_t>(p.kk + B);
        p.empty = ws.empty.get<int32_t>(p.kk);
        p.meta = ws.meta.get<int32_t>(2 * int64_t(B));
        p.qpre = ws.qpre.get<int64_t>(2 * (int64_t(B) + 1));
        p.d64pre = p.qpre + (B + 1);
        p.ticket = ws.ticket.get<uint32_t>(1);
        fill<uint32_t>(ctx, p.ticket, 1, 0u);
        if (p.csort) {
            p.bhist = ws.bhist.get<int32_t>(p.hh);
            allow_scatter_lds();
        }
        p.part = ws.part.get<double>(p.pp);
        if (!p.csort) {
            p.tmp_bytes = 0;
            rocprim::counting_iterator<int32_t> ids(0);
            KM_HIP(rocprim::radix_sort_pairs(nullptr, p.tmp_bytes, p.skeys, p.skeys + p.rows, ids, p.perm,
                                             unsigned(p.rows), 0, p.kbits, ctx->stream));
            p.tmp = ws.sort_tmp.get<char>(p.tmp_bytes);
        }
    }
    KM_HIP(hipMemcpyAsync(p.dI, p.hi.data(), sizeof(batch_inst) * B, hipMemcpyHostToDevice, ctx->stream));
    KM_HIP(hipMemcpyAsync(p.dpre, p.pre.data(), sizeof(int64_t) * p.pre.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    KM_HIP(hipMemcpyAsync(p.d_asel, p.asel.data(), sizeof(int32_t) * p.asel.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    KM_HIP(hipMemcpyAsync(p.d_apre, p.apre.data(), sizeof(int64_t) * p.apre.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    // the matrix-core kernels take more than 64 KB of LDS (set outside any capture)
    for (const batch_plan::asg_launch& L : p.launches) {
        if (L.v < BATCH_TILE_MFMA) continue;
        KM_HIP(hipFuncSetAttribute((const void*)batch_assign_mfma_kernel<32>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(batch_mfma_lds_bytes(32))));
        KM_HIP(hipFuncSetAttribute((const void*)batch_assign_mfma_kernel<64>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(batch_mfma_lds_bytes(64))));
        KM_HIP(hipFuncSetAttribute((const void*)batch_assign_mfma_kernel<128>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   int(batch_mfma_lds_bytes(128))));
        break;
    }
}

// exact assignment of every instance of the plan, one launch per tile variant in use
static void plan_assign(km_ctx* ctx, const batch_plan& p, int32_t* labels, uint32_t* skeys, int to_user) {
    for (const batch_plan::asg_launch& L : p.launches) {
        const dim3 grid(batch_grid(ctx, L.batches)), block(256);
        const int32_t* sel = p.d_asel + L.sel_off;
        const int64_t* pre = p.d_apre + L.pre_off;
        const int lb = int(p.lbits);
#define KM_BATCH_ASSIGN(CT, RW)                                                                      \
    hipLaunchKernelGGL((batch_assign_kernel<CT, RW>), grid, block, 0, ctx->stream, p.dI, sel, L.nsel, pre, \
                       p.Ccur, labels, skeys, lb, to_user)
#define KM_BATCH_MFMA(DP)                                                                          \
    hipLaunchKernelGGL((batch_assign_mfma_kernel<DP>), grid, block, batch_mfma_lds_bytes(DP), ctx->stream, \
                       p.dI, sel, L.nsel, pre, p.Ccur, labels, skeys, lb, to_user)
        switch (L.v) {
            case 0: KM_BATCH_ASSIGN(64, 8); break;
            case 1: KM_BATCH_ASSIGN(64, 4); break;
            case 2: KM_BATCH_ASSIGN(32, 8); break;
            case 3: KM_BATCH_ASSIGN(16, 8); break;
            case 4: KM_BATCH_ASSIGN(8, 8); break;
            case 5: KM_BATCH_ASSIGN(32, 4); break;
            case 6: KM_BATCH_ASSIGN(16, 4); break;
            case 7: KM_BATCH_MFMA(32); break;
            case 8: KM_BATCH_MFMA(64); break;
            case 9: KM_BATCH_MFMA(128); break;
            default: KM_BATCH_MFMA(16); break;
        }
#undef KM_BATCH_ASSIGN
#undef KM_BATCH_MFMA
        KM_LAUNCH_CHECK();
    }
}

// dir 0: every instance's C0 into the plan's concatenated centroids;
// dir 1: those centroids out to every instance's C_out (skipped where it is null)
static void plan_copy_centroids(km_ctx* ctx, const batch_plan& p, int dir) {
    hipLaunchKernelGGL(batch_centroid_copy_kernel,
                       dim3(batch_grid(ctx, (p.P(batch_plan::P_EL, p.B) + 255) / 256)), dim3(256), 0,
                       ctx->stream, p.dI, p.B, p.DP(batch_plan::P_EL), p.Ccur, dir);
    KM_LAUNCH_CHECK();
}

// one Lloyd iteration of every instance of a training plan; with events, the
// assignment stage is bracketed by them (eager runs only)
static void plan_iteration(km_ctx* ctx, batch_plan& p, hipEvent_t asg_start = nullptr,
                           hipEvent_t asg_end = nullptr) {
    using BP = batch_plan;
    const int B = p.B;
    const int cap = ctx->cu_count * 8;
    auto g = [&](int64_t items) { return batch_grid(ctx, items); };
    if (asg_start) KM_HIP(hipEventRecord(asg_start, ctx->stream));
    plan_assign(ctx, p, p.labels, p.csort ? nullptr : p.skeys, 0);
    if (asg_end) KM_HIP(hipEventRecord(asg_end, ctx->stream));
    if (!p.csort) fill<int32_t>(ctx, p.counts, p.kk, 0);
    // counting sort: per-block histograms only (the scan sums them into the counts)
    hipLaunchKernelGGL(batch_count_kernel, dim3(g(p.P(BP::P_CNT, B))), dim3(256), 0, ctx->stream, p.dI, B,
                       p.DP(BP::P_CNT), p.labels, p.csort ? (int32_t*)nullptr : p.counts,
                       p.csort ? p.bhist : (int32_t*)nullptr);
    KM_LAUNCH_CHECK();
    hipLaunchKernelGGL(batch_scan_kernel, dim3(g(B)), dim3(1024), 0, ctx->stream, p.dI, B, p.counts,
                       p.csort ? p.bhist : (int32_t*)nullptr, p.offsets, p.cb, p.empty, p.meta, p.qpre,
                       p.d64pre, p.ticket);
    KM_LAUNCH_CHECK();
    if (p.csort) {
        hipLaunchKernelGGL(batch_scatter_kernel, dim3(g(p.P(BP::P_CNT, B))), dim3(256),
                           batch_scatter_lds_bytes(p.maxk),
                           ctx->stream, p.dI, B, p.DP(BP::P_CNT), p.labels, p.bhist, p.perm, p.maxk);
        KM_LAUNCH_CHECK();
    } else {
        rocprim::counting_iterator<int32_t> ids(0);
        size_t tmp_bytes = p.tmp_bytes;
        KM_HIP(rocprim::radix_sort_pairs(p.tmp, tmp_bytes, p.skeys, p.skeys + p.rows, ids, p.perm,
                                         unsigned(p.rows), 0, p.kbits, ctx->stream));
    }
    if (p.any_vec) {
        hipLaunchKernelGGL(batch_reduce_vec_kernel, dim3(cap), dim3(256), 0, ctx->stream, p.dI, B, p.qpre,
                           p.perm, p.offsets, p.cb, p.part);
        KM_LAUNCH_CHECK();
    }
    if (p.any_scalar) {
        hipLaunchKernelGGL(batch_reduce_kernel, dim3(cap), dim3(256), 0, ctx->stream, p.dI, B, p.qpre,
                           p.perm, p.offsets, p.cb, p.part);
        KM_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(batch_finalize_kernel, dim3(g((p.P(BP::P_EL, B) + 255) / 256)), dim3(256), 0,
                       ctx->stream, p.dI, B, p.DP(BP::P_EL), p.counts, p.cb, p.part, p.Cnext);
    KM_LAUNCH_CHECK();
    // blocks of the instances with empty clusters only (the scan's prefix)
    hipLaunchKernelGGL(batch_d64_kernel, dim3(g(p.P(BP::P_D64, B))), dim3(256), 0, ctx->stream, p.dI, B,
                       p.d64pre, p.meta, p.Ccur, p.labels, p.d64);
    KM_LAUNCH_CHECK();
    hipLaunchKernelGGL(batch_fill_empty_kernel, dim3(g(B)), dim3(256), 0, ctx->stream, p.dI, B, p.meta,
                       p.empty, p.d64, p.Ccur, p.Cnext);
    KM_LAUNCH_CHECK();
    // every centroid of Cnext is written (finalize: counted clusters; fill:
    // empty ones): the next iteration reads it, and this one's input becomes
    // the next output buffer
    std::swap(p.Ccur, p.Cnext);
}

// final pass: labels into every instance's labels_out, objectives into p.objd,
// centroids out
static void plan_finish(km_ctx* ctx, const batch_plan& p) {
    using BP = batch_plan;
    const int B = p.B;
    auto g = [&](int64_t items) { return batch_grid(ctx, items); };
    plan_assign(ctx, p, nullptr, nullptr, 1);
    // every row's d64 (against its final label), then the chunk sums in order
    hipLaunchKernelGGL(batch_d64_kernel, dim3(g(p.P(BP::P_D64, B))), dim3(256), 0, ctx->stream, p.dI, B,
                       p.DP(BP::P_D64), (const int32_t*)nullptr, p.Ccur, (const int32_t*)nullptr, p.d64);
    KM_LAUNCH_CHECK();
    hipLaunchKernelGGL(batch_objective_kernel, dim3(g(p.P(BP::P_OBJ, B))), dim3(256), 0, ctx->stream, p.dI,
                       B, p.DP(BP::P_OBJ), p.d64, p.objc);
    KM_LAUNCH_CHECK();
    hipLaunchKernelGGL(batch_objective_sum_kernel, dim3((B + 255) / 256), dim3(256), 0, ctx->stream, p.dI,
                       B, p.objc, p.objd);
    KM_LAUNCH_CHECK();
    plan_copy_centroids(ctx, p, 1);
}

static void batch_train(km_ctx* ctx, const km_batch_instance_f32* inst, int64_t count,
                        int64_t max_iter, uint32_t flags, double* obj_out, int64_t* iters_out) {
    if (!inst || count <= 0) throw status_error(KM_DOMAIN_ERROR, "Input batch is empty");
    if (max_iter < 0)
        throw status_error(KM_DOMAIN_ERROR, "Max iteration count should be >= 0");
    if (!obj_out || !iters_out) throw status_error(KM_INVALID_ARGUMENT, "Output buffers are required");
    batch_plan p;
    plan_build(p, ctx, inst, count, true);
    ctx->stats = km_stats{};
    plan_upload(ctx, ctx->bws, p);
    KM_HIP(hipStreamSynchronize(ctx->stream));  // the host tables may be released from here on
    plan_copy_centroids(ctx, p, 0);

    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->timing) {
        ctx->ev_used = 0;
        e0 = ctx->event();
        e1 = ctx->event();
        KM_HIP(hipEventRecord(e0, ctx->stream));
    }
    if ((flags & KM_FLAG_GRAPH) && max_iter > 0) {
        // the whole job -- every iteration and the final pass -- is one graph,
        // launched once
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        KM_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        try {
            for (int64_t it = 0; it < max_iter; ++it) plan_iteration(ctx, p);
            plan_finish(ctx, p);
        } catch (...) {
            hipGraph_t broken = nullptr;
            (void)hipStreamEndCapture(ctx->stream, &broken);
            if (broken) (void)hipGraphDestroy(broken);
            throw;
        }
        KM_HIP(hipStreamEndCapture(ctx->stream, &graph));
        hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            (void)hipGraphDestroy(graph);
            KM_HIP(ie);
        }
        const hipError_t le = hipGraphLaunch(exec, ctx->stream);
        // the executable graph may be destroyed while its launch is in flight
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        KM_HIP(le);
        if (ctx->timing) KM_HIP(hipEventRecord(e1, ctx->stream));
        KM_HIP(hipMemcpyAsync(obj_out, p.objd, sizeof(double) * p.B, hipMemcpyDeviceToHost, ctx->stream));
        KM_HIP(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < p.B; ++i) iters_out[i] = max_iter;
        if (ctx->timing) {
            float ms = 0.f;
            KM_HIP(hipEventElapsedTime(&ms, e0, e1));
            ctx->stats.iterations_ms = ms;
        }
        return;
    } else {
        // eager, timed: events around every iteration's assignment stage
        std::vector<hipEvent_t> asg;
        for (int64_t it = 0; it < max_iter; ++it) {
            hipEvent_t a = nullptr, b = nullptr;
            if (ctx->timing) {
                a = ctx->event();
                b = ctx->event();
                asg.push_back(a);
                asg.push_back(b);
            }
            plan_iteration(ctx, p, a, b);
        }
        if (ctx->timing) {
            // the events are read after the synchronisation below
            ctx->stats.assign_launches = int64_t(asg.size() / 2);
        }
        plan_finish(ctx, p);
        if (ctx->timing) KM_HIP(hipEventRecord(e1, ctx->stream));
        KM_HIP(hipMemcpyAsync(obj_out, p.objd, sizeof(double) * p.B, hipMemcpyDeviceToHost, ctx->stream));
        KM_HIP(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < p.B; ++i) iters_out[i] = max_iter;
        if (ctx->timing) {
            float ms = 0.f;
            KM_HIP(hipEventElapsedTime(&ms, e0, e1));
            ctx->stats.iterations_ms = ms;
            double sum = 0.0;
            for (size_t i = 0; i + 1 < asg.size(); i += 2) {
                KM_HIP(hipEventElapsedTime(&ms, asg[i], asg[i + 1]));
                sum += ms;
            }
            ctx->stats.assign_kernel_ms = sum;
        }
        return;
    }
    // final assignment (into the callers' label buffers), objective, centroids out
    plan_finish(ctx, p);
    if (ctx->timing) KM_HIP(hipEventRecord(e1, ctx->stream));
    KM_HIP(hipMemcpyAsync(obj_out, p.objd, sizeof(double) * p.B, hipMemcpyDeviceToHost, ctx->stream));
    KM_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < p.B; ++i) iters_out[i] = max_iter;
    if (ctx->timing) {
        float ms = 0.f;
        KM_HIP(hipEventElapsedTime(&ms, e0, e1));
        ctx->stats.iterations_ms = ms;
    }
}

// One step of a continuous batch (km_batch_step_f32): `iterations` Lloyd
// iterations of every instance of `it` (C0 in, C_out out; they may alias), then
// a final pass over every instance of `fin` (labels, objective, centroids out
// where C_out is set).  The final pass runs after the iterations, so it sees
// centroids the iterated set of the same step wrote.
static void batch_step(km_ctx* ctx, const km_batch_instance_f32* it, int64_t n_it, int64_t iterations,
                       const km_batch_instance_f32* fin, int64_t n_fin, uint32_t flags,
                       double* obj_out) {
    if (n_it < 0 || n_fin < 0 || (n_it == 0 && n_fin == 0))
        throw status_error(KM_DOMAIN_ERROR, "Input batch is empty");
    if (n_it > 0 && iterations < 0)
        throw status_error(KM_DOMAIN_ERROR, "Max iteration count should be >= 0");
    if (n_fin > 0 && !obj_out) throw status_error(KM_INVALID_ARGUMENT, "Output buffers are required");
    if (flags != 0) throw status_error(KM_INVALID_ARGUMENT, "no flags are defined for a batched step");
    batch_plan pi, pf;
    if (n_it > 0) plan_build(pi, ctx, it, n_it, true);
    if (n_fin > 0) plan_build(pf, ctx, fin, n_fin, false);
    ctx->stats = km_stats{};
    if (n_it > 0 && iterations > 0) {
        plan_upload(ctx, ctx->bws, pi);
        plan_copy_centroids(ctx, pi, 0);
        for (int64_t t = 0; t < iterations; ++t) plan_iteration(ctx, pi);
        plan_copy_centroids(ctx, pi, 1);
    } else if (n_it > 0) {
        // zero iterations: centroids are carried over unchanged
        plan_upload(ctx, ctx->bws, pi);
        plan_copy_centroids(ctx, pi, 0);
        plan_copy_centroids(ctx, pi, 1);
    }
    if (n_fin > 0) {
        plan_upload(ctx, ctx->bws_fin, pf);
        plan_copy_centroids(ctx, pf, 0);
        plan_finish(ctx, pf);
        KM_HIP(hipMemcpyAsync(obj_out, pf.objd, sizeof(double) * pf.B, hipMemcpyDeviceToHost,
                              ctx->stream));
    }
    // the plans' host tables are read by the stream until here
    KM_HIP(hipStreamSynchronize(ctx->stream));
}

// Tile autotuner (km_batch_autotune): for every (d, k) bucket, synthetic
// clustered instances of the bucket's largest shape (rows near one of k random
// centroids) are assigned with every tile variant;
// the fastest (HIP-event time, best of three after a warm-up launch) becomes the
// context's variant for the bucket.  Every variant computes the same labels.
static void batch_autotune(km_ctx* ctx, int32_t* table_out, int64_t table_len) {
    static const int dreps[7] = {8, 16, 32, 64, 128, 256, 512};
    static const int kreps[6] = {8, 16, 32, 64, 128, 256};
    constexpr int NI = 16;  // instances per timed launch
    struct force_guard {
        km_ctx* c;
        int f;
        ~force_guard() { c->tile_force = f; }
    } guard{ctx, ctx->tile_force};
    int8_t table[KM_BATCH_TILE_BUCKETS];
    ctx->ev_used = 0;
    hipEvent_t e0 = ctx->event(), e1 = ctx->event();
    for (int db = 0; db < 7; ++db) {
        for (int kb = 0; kb < 6; ++kb) {
            const int d = dreps[db], k = kreps[kb];
            // about 1e11 distance operations per launch, X at most 1 GiB
            int64_t rows = int64_t(1e11 / (3.0 * k * d));
            rows = std::min<int64_t>(std::max<int64_t>(rows, int64_t(1) << 18), int64_t(1) << 22);
            rows = std::min<int64_t>(rows, (int64_t(1) << 28) / d);
            const int64_t n = rows / NI;
            float* X = ctx->tune_x.get<float>(NI * n * d);
            float* C = ctx->tune_c.get<float>(int64_t(NI) * k * d);
            int32_t* lab = ctx->tune_l.get<int32_t>(NI * n);
            hipLaunchKernelGGL(batch_tune_fill_kernel, dim3(batch_grid(ctx, (int64_t(NI) * k * d + 255) / 256)),
                               dim3(256), 0, ctx->stream, C, int64_t(NI) * k * d, 0x9876u + uint32_t(db * 6 + kb));
            KM_LAUNCH_CHECK();
            hipLaunchKernelGGL(batch_tune_points_kernel, dim3(batch_grid(ctx, (NI * n * d + 255) / 256)),
                               dim3(256), 0, ctx->stream, X, C, NI, n, d, k, 0x1234u + uint32_t(db * 6 + kb));
            KM_LAUNCH_CHECK();
            std::vector<km_batch_instance_f32> q(NI);
            for (int i = 0; i < NI; ++i)
                q[i] = km_batch_instance_f32{X + i * n * d, C + int64_t(i) * k * d, nullptr, lab + i * n, n, d, k};
            float best_ms = std::numeric_limits<float>::infinity();
            int best_v = default_tile(k);
            for (int v = 0; v < BATCH_TILES; ++v) {
                if (v == BATCH_TILE_MFMA && !mfma_tile_ok(d, k)) continue;
                ctx->tile_force = v + 1;
                batch_plan p;
                plan_build(p, ctx, q.data(), NI, false);
                plan_upload(ctx, ctx->bws_fin, p);
                plan_copy_centroids(ctx, p, 0);
                plan_assign(ctx, p, nullptr, nullptr, 1);  // warm-up
                float ms_v = std::numeric_limits<float>::infinity();
                for (int rep = 0; rep < 3; ++rep) {
                    KM_HIP(hipEventRecord(e0, ctx->stream));
                    plan_assign(ctx, p, nullptr, nullptr, 1);
                    KM_HIP(hipEventRecord(e1, ctx->stream));
                    KM_HIP(hipEventSynchronize(e1));
                    float ms = 0.f;
                    KM_HIP(hipEventElapsedTime(&ms, e0, e1));
                    ms_v = std::min(ms_v, ms);
                }
                KM_HIP(hipStreamSynchronize(ctx->stream));  // the plan's host tables go out of scope
                if (ms_v < best_ms) {
                    best_ms = ms_v;
                    best_v = v;
                }
            }
            table[db * 6 + kb] = int8_t(best_v);
        }
    }
    std::memcpy(ctx->tile_table, table, sizeof(table));
    ctx->tune_x.release();
    ctx->tune_c.release();
    ctx->tune_l.release();
    if (table_out)
        for (int64_t i = 0; i < std::min<int64_t>(table_len, KM_BATCH_TILE_BUCKETS); ++i) table_out[i] = table[i];
}

// ---------------------------------------------------------------------------
// Initial centroids (km_init_centroids_*): kmeans_init's dense, random_dense and
// plus_plus_dense methods.  Preconditions of kmeans_init/detail/compute_ops.hpp:
// 40-56 (checked in that order).
// ---------------------------------------------------------------------------
// k distinct row ids of [0, n): a partial Fisher-Yates shuffle driven by the
// counter-based generator (random_dense)
static std::vector<int64_t> init_random_rows(uint64_t seed, int64_t n, int64_t k) {
    std::vector<int64_t> ids(k);
    std::unordered_map<int64_t, int64_t> moved;
    for (int64_t i = 0; i < k; ++i) {
        const int64_t j = i + int64_t(rng_index(seed, RNG_STREAM_RANDOM, uint64_t(i), uint64_t(n - i)));
        auto at = [&](int64_t p) {
            auto it = moved.find(p);
            return it == moved.end() ? p : it->second;
        };
        const int64_t vi = at(i), vj = at(j);
        ids[i] = vj;
        moved[j] = vi;
    }
    return ids;
}

static void init_check_args(const void* X, int64_t n, int64_t d, int64_t k, int method, const void* C_out) {
    if (X == nullptr || n <= 0 || d <= 0) throw status_error(KM_DOMAIN_ERROR, "Input data is empty");
    if (k <= 0) throw status_error(KM_DOMAIN_ERROR, "Cluster count is lower than or equal to zero");
    if (k >= n) throw status_error(KM_DOMAIN_ERROR, "Cluster count exceeds data row count");
    if (C_out == nullptr) throw status_error(KM_INVALID_ARGUMENT, "Output buffers are required");
    if (method == KM_INIT_PARALLEL_PLUS_DENSE)
        throw status_error(KM_UNIMPLEMENTED, "K-Means init++ parallel dense method is not implemented for GPU");
    if (method < KM_INIT_DENSE || method > KM_INIT_PLUS_PLUS_DENSE)
        throw status_error(KM_INVALID_ARGUMENT, "unknown initialisation method");
    if (d > std::numeric_limits<int32_t>::max())
        throw status_error(KM_DOMAIN_ERROR, "Column count is greater than max int32 value");
}

template <typename F>
static void init_centroids(km_ctx* ctx, const F* X, int64_t n, int64_t d, int64_t k, int method,
                           uint64_t seed, uint32_t flags, F* C_out) {
    init_check_args(X, n, d, k, method, C_out);
    const bool dev = (flags & KM_FLAG_DEVICE_POINTERS) != 0;
    const F* Xd = X;
    F* Cd = C_out;
    if (!dev) {
        // dense reads only the first k rows
        const int64_t up_rows = method == KM_INIT_DENSE ? k : n;
        F* xb = ctx->x_dev.get<F>(up_rows * d);
        KM_HIP(hipMemcpyAsync(xb, X, sizeof(F) * up_rows * d, hipMemcpyHostToDevice, ctx->stream));
        Xd = xb;
        Cd = ctx->c_out_dev.get<F>(k * d);
    }
    const int di = int(d);
    if (method == KM_INIT_DENSE) {
        // the first k rows (compute_kernels_impl.hpp:53-68)
        KM_HIP(hipMemcpyAsync(Cd, Xd, sizeof(F) * k * d, hipMemcpyDeviceToDevice, ctx->stream));
    } else if (method == KM_INIT_RANDOM_DENSE) {
        // k distinct rows by a partial Fisher-Yates shuffle of [0, n) on the host,
        // gathered on the device (compute_kernel_distr_random_dense_dpc.cpp:33-66)
        const std::vector<int64_t> ids = init_random_rows(seed, n, k);
        int64_t* dids = ctx->init_ids.get<int64_t>(k);
        KM_HIP(hipMemcpyAsync(dids, ids.data(), sizeof(int64_t) * k, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(init_gather_rows_kernel<F>, dim3(grid_for(ctx, k, 1, 4)), dim3(64), 0, ctx->stream, Xd,
                           di, dids, k, Cd);
        KM_LAUNCH_CHECK();
        KM_HIP(hipStreamSynchronize(ctx->stream));  // ids is read by the copy above
    } else {
        // k-means++ (D^2 sampling), every step on the device without host syncs
        double* D2 = ctx->init_d2.get<double>(n);
        const int64_t nchunks = (n + OBJ_CHUNK - 1) / OBJ_CHUNK;
        double* chunks = ctx->init_chunks.get<double>(nchunks);
        int64_t* chosen = ctx->init_chosen.get<int64_t>(1);
        const int64_t first = int64_t(rng_index(seed, RNG_STREAM_PLUS_PLUS, 0, uint64_t(n)));
        hipLaunchKernelGGL(init_copy_row_kernel<F>, dim3(1), dim3(256), 0, ctx->stream, Xd, di, nullptr, first,
                           Cd, int64_t(0));
        KM_LAUNCH_CHECK();
        const int d2_grid = grid_for(ctx, n, 256, 8);
        const int ch_grid = grid_for(ctx, nchunks, 1, 8);
        for (int64_t t = 1; t < k; ++t) {
            hipLaunchKernelGGL(init_d2_kernel<F>, dim3(d2_grid), dim3(256), 0, ctx->stream, Xd, n, di,
                               Cd + (t - 1) * d, D2, t == 1 ? 1 : 0);
            KM_LAUNCH_CHECK();
            hipLaunchKernelGGL(init_chunk_sums_kernel, dim3(ch_grid), dim3(256), 0, ctx->stream, D2, n, chunks);
            KM_LAUNCH_CHECK();
            const double u = rng_uniform01(seed, RNG_STREAM_PLUS_PLUS, uint64_t(t));
            hipLaunchKernelGGL(init_select_kernel, dim3(1), dim3(256), 0, ctx->stream, D2, n, chunks, u, chosen);
            KM_LAUNCH_CHECK();
            hipLaunchKernelGGL(init_copy_row_kernel<F>, dim3(1), dim3(256), 0, ctx->stream, Xd, di, chosen,
                               int64_t(0), Cd, t);
            KM_LAUNCH_CHECK();
        }
    }
    if (!dev) KM_HIP(hipMemcpyAsync(C_out, Cd, sizeof(F) * k * d, hipMemcpyDeviceToHost, ctx->stream));
    KM_HIP(hipStreamSynchronize(ctx->stream));
}

// Rows of a CSR matrix with each row's column indices strictly increasing and
// zero-based: entries sorted by column (stable), the values of a repeated column
// added in stored order onto a zero, as the row's dense expansion holds them.
template <typename F>
static void csr_merge_rows(const F* vals, const int64_t* cols, const int64_t* offs, int64_t base, int64_t n,
                           std::vector<F>& mv, std::vector<int64_t>& mc, std::vector<int64_t>& mo) {
    mv.clear();
    mc.clear();
    mo.assign(size_t(n + 1), 0);
    std::vector<int64_t> idx;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t o0 = offs[i] - base, o1 = offs[i + 1] - base;
        idx.resize(size_t(o1 - o0));
        for (int64_t p = o0; p < o1; ++p) idx[size_t(p - o0)] = p;
        std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return cols[a] < cols[b]; });
        for (size_t j = 0; j < idx.size();) {
            const int64_t col = cols[idx[j]];
            F v = F(0);
            for (; j < idx.size() && cols[idx[j]] == col; ++j) v = v + vals[idx[j]];
            mc.push_back(col - base);
            mv.push_back(v);
        }
        mo[size_t(i + 1)] = int64_t(mc.size());
    }
}

// Initial centroids of CSR rows (km_init_centroids_csr_*): the dense methods
// applied to the rows' dense expansion, which is never formed.  The reference
// starts a lloyd_csr training without initial centroids with DAAL's k-means++
// on the host (train_kernel_lloyd_csr_dpc.cpp:41-68).  Every step is the dense
// driver's with the CSR row walk in place of the row read, so the centroids
// equal init_centroids on the expanded matrix bit for bit.
template <typename F>
static void init_centroids_csr(km_ctx* ctx, const F* vals, const int64_t* cols, const int64_t* offs,
                               int64_t n, int64_t d, int64_t nnz, int32_t indexing, int64_t k, int method,
                               uint64_t seed, uint32_t flags, F* C_out) {
    init_check_args(offs, n, d, k, method, C_out);
    const bool dev = (flags & KM_FLAG_DEVICE_POINTERS) != 0;
    bool increasing = true;
    csr_view<F> x = csr_prepare<F>(ctx, vals, cols, offs, n, d, nnz, indexing, dev, &increasing);
    // merged rows (only when a row has unsorted or repeated columns), kept alive
    // until the stream has consumed them
    std::vector<F> mv;
    std::vector<int64_t> mc, mo;
    if (!increasing) {
        std::vector<F> hv;
        std::vector<int64_t> hc, ho;
        const F* sv = vals;
        const int64_t* sc = cols;
        const int64_t* so = offs;
        if (dev) {
            hv.resize(size_t(nnz));
            hc.resize(size_t(nnz));
            ho.resize(size_t(n + 1));
            if (nnz > 0) {
                KM_HIP(hipMemcpyAsync(hv.data(), vals, sizeof(F) * nnz, hipMemcpyDeviceToHost, ctx->stream));
                KM_HIP(hipMemcpyAsync(hc.data(), cols, sizeof(int64_t) * nnz, hipMemcpyDeviceToHost, ctx->stream));
            }
            KM_HIP(hipMemcpyAsync(ho.data(), offs, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, ctx->stream));
            KM_HIP(hipStreamSynchronize(ctx->stream));
            sv = hv.data();
            sc = hc.data();
            so = ho.data();
        }
        csr_merge_rows<F>(sv, sc, so, x.base, n, mv, mc, mo);
        const int64_t mnz = int64_t(mc.size());
        F* vb = ctx->csr_vals.get<F>(std::max<int64_t>(mnz, 1));
        int64_t* cb = ctx->csr_cols.get<int64_t>(std::max<int64_t>(mnz, 1));
        int64_t* ob = ctx->csr_offs.get<int64_t>(n + 1);
        if (mnz > 0) {
            KM_HIP(hipMemcpyAsync(vb, mv.data(), sizeof(F) * mnz, hipMemcpyHostToDevice, ctx->stream));
            KM_HIP(hipMemcpyAsync(cb, mc.data(), sizeof(int64_t) * mnz, hipMemcpyHostToDevice, ctx->stream));
        }
        KM_HIP(hipMemcpyAsync(ob, mo.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, ctx->stream));
        KM_HIP(hipStreamSynchronize(ctx->stream));
        x = csr_view<F>{vb, cb, ob, 0, n, d, mnz};
    }
    const int di = int(d);
    F* Cd = dev ? C_out : ctx->c_out_dev.get<F>(k * d);
    std::vector<int64_t> ids;
    if (method == KM_INIT_DENSE || method == KM_INIT_RANDOM_DENSE) {
        const int64_t* dids = nullptr;
        if (method == KM_INIT_RANDOM_DENSE) {
            ids = init_random_rows(seed, n, k);
            int64_t* b = ctx->init_ids.get<int64_t>(k);
            KM_HIP(hipMemcpyAsync(b, ids.data(), sizeof(int64_t) * k, hipMemcpyHostToDevice, ctx->stream));
            dids = b;
        }
        hipLaunchKernelGGL(init_expand_csr_rows_kernel<F>, dim3(grid_for(ctx, k, 1, 4)), dim3(256), 0,
                           ctx->stream, x.vals, x.cols, x.offs, x.base, di, dids, nullptr, int64_t(0), k, Cd,
                           int64_t(0));
        KM_LAUNCH_CHECK();
        KM_HIP(hipStreamSynchronize(ctx->stream));  // ids is read by the expansion above
    } else {
        double* D2 = ctx->init_d2.get<double>(n);
        const int64_t nchunks = (n + OBJ_CHUNK - 1) / OBJ_CHUNK;
        double* chunks = ctx->init_chunks.get<double>(nchunks);
        int64_t* chosen = ctx->init_chosen.get<int64_t>(1);
        const int64_t first = int64_t(rng_index(seed, RNG_STREAM_PLUS_PLUS, 0, uint64_t(n)));
        hipLaunchKernelGGL(init_expand_csr_rows_kernel<F>, dim3(1), dim3(256), 0, ctx->stream, x.vals, x.cols,
                           x.offs, x.base, di, nullptr, nullptr, first, int64_t(1), Cd, int64_t(0));
        KM_LAUNCH_CHECK();
        const int d2_grid = grid_for(ctx, n, 256, 8);
        const int ch_grid = grid_for(ctx, nchunks, 1, 8);
        for (int64_t t = 1; t < k; ++t) {
            hipLaunchKernelGGL(init_d2_csr_kernel<F>, dim3(d2_grid), dim3(256), 0, ctx->stream, x.vals, x.cols,
                               x.offs, x.base, n, di, Cd + (t - 1) * d, D2, t == 1 ? 1 : 0);
            KM_LAUNCH_CHECK();
            hipLaunchKernelGGL(init_chunk_sums_kernel, dim3(ch_grid), dim3(256), 0, ctx->stream, D2, n, chunks);
            KM_LAUNCH_CHECK();
            const double u = rng_uniform01(seed, RNG_STREAM_PLUS_PLUS, uint64_t(t));
            hipLaunchKernelGGL(init_select_kernel, dim3(1), dim3(256), 0, ctx->stream, D2, n, chunks, u, chosen);
            KM_LAUNCH_CHECK();
            hipLaunchKernelGGL(init_expand_csr_rows_kernel<F>, dim3(1), dim3(256), 0, ctx->stream, x.vals,
                               x.cols, x.offs, x.base, di, nullptr, chosen, int64_t(0), int64_t(1), Cd, t);
            KM_LAUNCH_CHECK();
        }
    }
    if (!dev) KM_HIP(hipMemcpyAsync(C_out, Cd, sizeof(F) * k * d, hipMemcpyDeviceToHost, ctx->stream));
    KM_HIP(hipStreamSynchronize(ctx->stream));
}

// Device address of the scratch array on `device`, which must be the current
// device.  Looked up once per device (each device holds its own copy at its own
// address) and kept for the life of the process.
static std::mutex scratch_mu;
static void* scratch_dev[NOALLOC_MAX_DEVICES];
static hipError_t scratch_rc[NOALLOC_MAX_DEVICES];
static bool scratch_done[NOALLOC_MAX_DEVICES];

static void* scratch_address(int device, std::string* err) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) {
        *err = "device index " + std::to_string(device) + " is beyond the scratch table";
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_done[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) (void)hipGetLastError();
        scratch_rc[device] = e;
        scratch_dev[device] = e == hipSuccess ? p : nullptr;
        scratch_done[device] = true;
    }
    if (!scratch_dev[device])
        *err = std::string("hipGetSymbolAddress: ") + hipGetErrorString(scratch_rc[device]);
    return scratch_dev[device];
}

template <typename Fn>
static km_status guarded(km_ctx* ctx, Fn fn) {
    if (!ctx) return KM_INVALID_ARGUMENT;
    try {
        KM_HIP(hipSetDevice(ctx->device));
        fn();
        ctx->last_error.clear();
        return KM_OK;
    } catch (const status_error& e) {
        ctx->last_error = e.what();
        if (e.status == KM_INTERNAL_ERROR || e.status == KM_DEVICE_BAD_ALLOC) {
            // leave the stream idle so the context stays usable
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipGetLastError();
        }
        return e.status;
    } catch (const std::bad_alloc&) {
        ctx->last_error = "host allocation failed";
        return KM_HOST_BAD_ALLOC;
    } catch (const std::exception& e) {
        ctx->last_error = e.what();
        return KM_INTERNAL_ERROR;
    }
}

}  // namespace km

extern "C" {

km_status km_context_create(int device, km_ctx** out) {
    if (!out) return KM_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return KM_INTERNAL_ERROR;
    }
    if (device < 0 || device >= count) return KM_INVALID_ARGUMENT;
    km_ctx* ctx = new (std::nothrow) km_ctx();
    if (!ctx) return KM_HOST_BAD_ALLOC;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return KM_INTERNAL_ERROR;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->cu_count = prop.multiProcessorCount;
    void* pin = nullptr;
    if (hipHostMalloc(&pin, 64, hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return KM_HOST_BAD_ALLOC;
    }
    ctx->pinned = static_cast<int32_t*>(pin);
    // the device is current: resolve its copy of the no-allocation scratch array
    ctx->scratch = km::scratch_address(device, &ctx->scratch_err);
    *out = ctx;
    return KM_OK;
}

km_status km_context_destroy(km_ctx* ctx) {
    if (!ctx) return KM_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    km::dbuf* bufs[] = {&ctx->x_bf16, &ctx->enorm, &ctx->xmax, &ctx->c_bf16, &ctx->cnb, &ctx->params,
                        &ctx->coll_rows, &ctx->coll_T, &ctx->coll_count, &ctx->coll_ids, &ctx->x_dev, &ctx->x_pad, &ctx->c_cur, &ctx->c_next, &ctx->cn, &ctx->cpad, &ctx->cmax,
                        &ctx->xnorm, &ctx->labels, &ctx->lists_rows, &ctx->lists_cnt,
                        &ctx->lists_ids, &ctx->lists_full, &ctx->lists_counters, &ctx->counts,
                        &ctx->offsets, &ctx->chunk_base, &ctx->empty_list, &ctx->meta, &ctx->perm,
                        &ctx->keys_out, &ctx->sort_tmp, &ctx->part, &ctx->d64, &ctx->obj_chunks,
                        &ctx->obj_dev, &ctx->hist, &ctx->cand_gt, &ctx->cand_eq,
                        &ctx->cand_dev, &ctx->c_out_dev, &ctx->seg_cand_rows,
                        &ctx->seg_cand_cnt, &ctx->seg_cand_ids, &ctx->seg_coll_rows, &ctx->seg_coll_T,
                        &ctx->seg_counts, &ctx->tune_x, &ctx->tune_c, &ctx->tune_l,
                        &ctx->init_d2, &ctx->init_chunks, &ctx->init_chosen, &ctx->init_ids,
                        &ctx->gt_d, &ctx->stat_acc, &ctx->coop_wcnt, &ctx->cs_inst, &ctx->cs_pre,
                        &ctx->cs_hist, &ctx->cs_gsum, &ctx->csr_vals, &ctx->csr_cols, &ctx->csr_offs, &ctx->csr_xs,
                        &ctx->csr_ct, &ctx->csr_cs, &ctx->csr_closest, &ctx->csr_flag};
    for (auto* b : bufs) b->release();
    ctx->bws.release();
    ctx->bws_fin.release();
    for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return KM_OK;
}

const char* km_last_error(const km_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

km_status km_context_set_option(km_ctx* ctx, int option, int64_t value) {
    if (!ctx) return KM_INVALID_ARGUMENT;
    switch (option) {
        case KM_OPT_ASSIGN_PATH:
            if (value < 0 || value > 2) return KM_INVALID_ARGUMENT;
            ctx->assign_path_opt = int(value);
            return KM_OK;
        case KM_OPT_TIMING:
            ctx->timing = value != 0;
            return KM_OK;
        case KM_OPT_BATCH_TILE:
            if (value < 0 || value > km::BATCH_TILES) return KM_INVALID_ARGUMENT;
            ctx->tile_force = int(value);
            return KM_OK;
        default:
            return KM_INVALID_ARGUMENT;
    }
}

km_status km_get_stats(const km_ctx* ctx, km_stats* out) {
    if (!ctx || !out) return KM_INVALID_ARGUMENT;
    *out = ctx->stats;
    return KM_OK;
}

km_status km_measure_bf16_mfma_peak(km_ctx* ctx, int64_t iters, double* tflops_out,
                                    double* clock_ghz_out) {
    return km::guarded(ctx, [&] {
        if (iters <= 0 || iters > (int64_t(1) << 24) || !tflops_out)
            throw km::status_error(KM_INVALID_ARGUMENT, "iterations must lie in [1, 2^24]");
        const int grid = ctx->cu_count * 2;
        // operands: small bf16 values (exponent near 1), all finite
        std::vector<uint32_t> h(2 * km::PROBE_WG_THREADS * 4);
        for (size_t i = 0; i < h.size(); ++i) h[i] = 0x3F803F80u ^ ((uint32_t(i) * 2654435761u) & 0x00070007u);
        uint32_t* src = ctx->tune_x.get<uint32_t>(h.size());
        float* out = ctx->tune_c.get<float>(int64_t(grid) * km::PROBE_WG_THREADS);
        long long* clk = ctx->tune_l.get<long long>(2);
        KM_HIP(hipMemcpyAsync(src, h.data(), h.size() * 4, hipMemcpyHostToDevice, ctx->stream));
        ctx->ev_used = 0;
        hipEvent_t e0 = ctx->event(), e1 = ctx->event();
        auto launch = [&] {
            hipLaunchKernelGGL(km::mfma_ceiling_kernel, dim3(grid), dim3(km::PROBE_WG_THREADS), 0,
                               ctx->stream, reinterpret_cast<const km::uintx4*>(src), out, clk, int(iters));
            KM_LAUNCH_CHECK();
        };
        launch();  // warm-up (clock ramp)
        KM_HIP(hipEventRecord(e0, ctx->stream));
        launch();
        KM_HIP(hipEventRecord(e1, ctx->stream));
        long long hc[2] = {0, 0};
        KM_HIP(hipMemcpyAsync(hc, clk, sizeof(hc), hipMemcpyDeviceToHost, ctx->stream));
        KM_HIP(hipStreamSynchronize(ctx->stream));
        float ms = 0.f;
        KM_HIP(hipEventElapsedTime(&ms, e0, e1));
        const double flops = double(grid) * (km::PROBE_WG_THREADS / 64) * 4.0 * double(iters) * 32 * 32 * 16 * 2;
        *tflops_out = flops / (double(ms) * 1e-3) / 1e12;
        if (clock_ghz_out) *clock_ghz_out = hc[1] > 0 ? double(hc[0]) / (double(hc[1]) / 100e6) / 1e9 : 0.0;
        ctx->tune_x.release();
        ctx->tune_c.release();
        ctx->tune_l.release();
    });
}

km_status km_lloyd_train_f32(km_ctx* ctx, const float* X, int64_t n, int64_t d, const float* C0,
                             int64_t k, int64_t max_iter, double acc_thr, uint32_t flags,
                             float* C_out, int32_t* labels_out, double* obj_out,
                             int64_t* iters_out) {
    return km::guarded(ctx, [&] {
        km::check_train_args(X, n, d, C0, k, max_iter, acc_thr, C_out, labels_out);
        km::train<float>(ctx, X, n, d, C0, k, max_iter, acc_thr, flags, C_out, labels_out, obj_out,
                         iters_out);
    });
}

km_status km_lloyd_train_f64(km_ctx* ctx, const double* X, int64_t n, int64_t d, const double* C0,
                             int64_t k, int64_t max_iter, double acc_thr, uint32_t flags,
                             double* C_out, int32_t* labels_out, double* obj_out,
                             int64_t* iters_out) {
    return km::guarded(ctx, [&] {
        km::check_train_args(X, n, d, C0, k, max_iter, acc_thr, C_out, labels_out);
        km::train<double>(ctx, X, n, d, C0, k, max_iter, acc_thr, flags, C_out, labels_out,
                          obj_out, iters_out);
    });
}

km_status km_lloyd_infer_f32(km_ctx* ctx, const float* X, int64_t n, int64_t d, const float* C,
                             int64_t k, uint32_t flags, int32_t* labels_out, double* obj_out) {
    return km::guarded(ctx, [&] {
        km::check_infer_args(X, n, d, C, k, flags, labels_out, obj_out);
        km::infer<float>(ctx, X, n, d, C, k, flags, labels_out, obj_out);
    });
}

km_status km_lloyd_infer_f32_noalloc(km_ctx* ctx, const float* X, int64_t n, int64_t d,
                                     const float* C, int64_t k, uint32_t flags, int32_t* labels_out,
                                     double* obj_out) {
    return km::guarded(ctx, [&] {
        km::check_infer_args(X, n, d, C, k, flags, labels_out, obj_out);
        km::infer_noalloc(ctx, X, n, d, C, k, flags, labels_out, obj_out);
    });
}

km_status km_lloyd_infer_f64(km_ctx* ctx, const double* X, int64_t n, int64_t d, const double* C,
                             int64_t k, uint32_t flags, int32_t* labels_out, double* obj_out) {
    return km::guarded(ctx, [&] {
        km::check_infer_args(X, n, d, C, k, flags, labels_out, obj_out);
        km::infer<double>(ctx, X, n, d, C, k, flags, labels_out, obj_out);
    });
}

#define KM_CSR_ABI(SUF, F)                                                                          \
    km_status km_lloyd_train_csr_##SUF(km_ctx* ctx, const F* values, const int64_t* col_indices,     \
                                       const int64_t* row_offsets, int64_t n, int64_t d,           \
                                       int64_t nnz, int32_t indexing, const F* C0, int64_t k,      \
                                       int64_t max_iter, double acc_thr, uint32_t flags,           \
                                       F* C_out, int32_t* labels_out, double* obj_out,             \
                                       int64_t* iters_out) {                                       \
        return km::guarded(ctx, [&] {                                                               \
            km::check_csr_args(values, col_indices, row_offsets, nnz, indexing);                    \
            km::check_train_args(row_offsets, n, d, C0, k, max_iter, acc_thr, C_out, labels_out);   \
            km::train_csr<F>(ctx, values, col_indices, row_offsets, n, d, nnz, indexing, C0, k,     \
                             max_iter, acc_thr, flags, C_out, labels_out, obj_out, iters_out);      \
        });                                                                                         \
    }                                                                                               \
    km_status km_lloyd_infer_csr_##SUF(km_ctx* ctx, const F* values, const int64_t* col_indices,     \
                                       const int64_t* row_offsets, int64_t n, int64_t d,           \
                                       int64_t nnz, int32_t indexing, const F* C, int64_t k,       \
                                       uint32_t flags, int32_t* labels_out, double* obj_out) {     \
        return km::guarded(ctx, [&] {                                                               \
            km::check_csr_args(values, col_indices, row_offsets, nnz, indexing);                    \
            km::check_infer_args(row_offsets, n, d, C, k, flags, labels_out, obj_out);              \
            km::infer_csr<F>(ctx, values, col_indices, row_offsets, n, d, nnz, indexing, C, k,      \
                             flags, labels_out, obj_out);                                           \
        });                                                                                         \
    }
KM_CSR_ABI(f32, float)
KM_CSR_ABI(f64, double)
#undef KM_CSR_ABI

km_status km_batch_train_f32(km_ctx* ctx, const km_batch_instance_f32* instances, int64_t count,
                             int64_t max_iter, uint32_t flags, double* obj_out,
                             int64_t* iters_out) {
    return km::guarded(ctx, [&] {
        km::batch_train(ctx, instances, count, max_iter, flags, obj_out, iters_out);
    });
}

km_status km_device_alloc(km_ctx* ctx, int64_t bytes, void** out) {
    return km::guarded(ctx, [&] {
        if (!out || bytes < 0) throw km::status_error(KM_INVALID_ARGUMENT, "allocation size and output are required");
        *out = nullptr;
        hipError_t e = hipMalloc(out, size_t(std::max<int64_t>(bytes, 1)));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *out = nullptr;
            throw km::status_error(KM_DEVICE_BAD_ALLOC,
                                   "device allocation of " + std::to_string(bytes) + " bytes failed");
        }
    });
}

km_status km_device_free(km_ctx* ctx, void* ptr) {
    return km::guarded(ctx, [&] {
        if (ptr) KM_HIP(hipFree(ptr));
    });
}

km_status km_memcpy(km_ctx* ctx, void* dst, const void* src, int64_t bytes, int32_t kind) {
    return km::guarded(ctx, [&] {
        if (bytes < 0 || (bytes > 0 && (!dst || !src)))
            throw km::status_error(KM_INVALID_ARGUMENT, "copy needs a size and both pointers");
        hipMemcpyKind k;
        switch (kind) {
            case KM_COPY_HOST_TO_DEVICE: k = hipMemcpyHostToDevice; break;
            case KM_COPY_DEVICE_TO_HOST: k = hipMemcpyDeviceToHost; break;
            case KM_COPY_DEVICE_TO_DEVICE: k = hipMemcpyDeviceToDevice; break;
            default: throw km::status_error(KM_INVALID_ARGUMENT, "unknown copy kind");
        }
        if (bytes == 0) return;
        KM_HIP(hipMemcpyAsync(dst, src, size_t(bytes), k, ctx->stream));
        KM_HIP(hipStreamSynchronize(ctx->stream));
    });
}

km_status km_init_centroids_f32(km_ctx* ctx, const float* X, int64_t n, int64_t d, int64_t k,
                                int32_t method, uint64_t seed, uint32_t flags, float* C_out) {
    return km::guarded(ctx, [&] { km::init_centroids<float>(ctx, X, n, d, k, method, seed, flags, C_out); });
}

km_status km_init_centroids_f64(km_ctx* ctx, const double* X, int64_t n, int64_t d, int64_t k,
                                int32_t method, uint64_t seed, uint32_t flags, double* C_out) {
    return km::guarded(ctx, [&] { km::init_centroids<double>(ctx, X, n, d, k, method, seed, flags, C_out); });
}

#define KM_INIT_CSR_ABI(SUF, F)                                                                       \
    km_status km_init_centroids_csr_##SUF(km_ctx* ctx, const F* values, const int64_t* col_indices,      \
                                          const int64_t* row_offsets, int64_t n, int64_t d, int64_t nnz, \
                                          int32_t indexing, int64_t k, int32_t method, uint64_t seed,    \
                                          uint32_t flags, F* C_out) {                                    \
        return km::guarded(ctx, [&] {                                                                     \
            km::check_csr_args(values, col_indices, row_offsets, nnz, indexing);                          \
            km::init_centroids_csr<F>(ctx, values, col_indices, row_offsets, n, d, nnz, indexing, k,      \
                                      method, seed, flags, C_out);                                        \
        });                                                                                               \
    }
KM_INIT_CSR_ABI(f32, float)
KM_INIT_CSR_ABI(f64, double)
#undef KM_INIT_CSR_ABI

km_status km_batch_autotune(km_ctx* ctx, int32_t* table_out, int64_t table_len) {
    return km::guarded(ctx, [&] {
        if (table_len < 0) throw km::status_error(KM_INVALID_ARGUMENT, "table length is negative");
        km::batch_autotune(ctx, table_out, table_len);
    });
}

km_status km_batch_step_f32(km_ctx* ctx, const km_batch_instance_f32* iterated, int64_t n_iterated,
                            int64_t iterations, const km_batch_instance_f32* finished,
                            int64_t n_finished, uint32_t flags, double* obj_out) {
    return km::guarded(ctx, [&] {
        km::batch_step(ctx, iterated, n_iterated, iterations, finished, n_finished, flags, obj_out);
    });
}

}  // extern "C"
