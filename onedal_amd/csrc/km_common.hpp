// km_common.hpp -- shared definitions of the MI355X k-means Lloyd engine.
//
// The engine replaces oneDAL's dense Lloyd GPU backend
// (cpp/oneapi/dal/algo/kmeans/backend/gpu/*) with hand-written CDNA4 kernels.
// Arithmetic contract (every path -- FP32 MFMA, BF16 MFMA + refine, FP64 --
// produces exactly these results):
//   * label_i = argmin_j D(x_i, c_j), D = sum_f (x_f - c_jf)^2 in the input
//     precision F, features in order, each sub/mul/add rounded on its own, ties to
//     the lowest j.  Fast paths compute an approximate distance with a proven
//     error bound and re-score in this exact form every row whose runner-up lies
//     inside the bound.
//   * centroid sums in double, in a fixed order (REDUCE_CHUNK, reduce_groups);
//     C_j = (F)(S1_j / count_j).
//   * objective = sum_i d64_i in a fixed order (objective_chunks_kernel), d64 =
//     sum_f ((double)x_f - (double)c_f)^2.
// All code here is compiled with -ffp-contract=off: fused multiply-adds appear
// only where fma() is written out.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace km {

// Rows of one cluster's member list summed by one wave before the partials are
// combined (fixed, so the FP64 sums are reproducible run to run).
constexpr int REDUCE_CHUNK = 256;
// Rows per objective partial sum.
constexpr int OBJ_CHUNK = 4096;
// Candidate ids carried per row to the exact re-score.
constexpr int MAXC = 8;

// Interleaved partial sums per reduction chunk: a wave of 64 lanes covers
// 64/d rows at a time when d <= 32.
__host__ __device__ inline int reduce_groups(int d) { return d <= 32 ? 64 / d : 1; }

template <typename F>
struct fp_traits;

template <>
struct fp_traits<float> {
    // unit roundoff of the accumulation format, doubled to cover accumulation
    // that is not round-to-nearest (matrix-core internal adds)
    static constexpr double u = 5.9604644775390625e-08;        // 2^-24
    static constexpr double u_acc = 1.1920928955078125e-07;    // 2^-23
    static constexpr float inf = __builtin_inff();
};

template <>
struct fp_traits<double> {
    static constexpr double u = 1.1102230246251565e-16;        // 2^-53
    static constexpr double u_acc = 2.220446049250313e-16;     // 2^-52
    static constexpr double inf = __builtin_inf();
};

// Candidate ids gathered per row by the BF16 collect pass.
constexpr int MAXC2 = 32;

// Device-side work lists filled by the assignment kernels.
struct refine_lists {
    int32_t* cand_rows;    // rows with 2..MAXC certified candidates
    int32_t* cand_count;   // number of candidates per entry
    int32_t* cand_ids;     // MAXC ids per entry
    int32_t* full_rows;    // rows re-scored against every centroid
    int32_t* counters;     // [0] #cand entries, [1] #full entries, [2] #collect entries
    // BF16 path: rows whose slots could not certify the candidate set
    int32_t* coll_rows;    // row ids
    double* coll_T;        // their thresholds (shifted value domain)
    int32_t* coll_count;   // ids gathered per entry (may exceed MAXC2: overflow)
    int32_t* coll_ids;     // MAXC2 ids per entry
};

// Per-iteration centroid statistics used by the error bound.
struct centroid_stats {
    double cmax;     // max_j ||c_j|| (rounded up)
};

// Error-bound constants for one assignment launch (host computed).
//   |v_j - V_j| <= da = a_xc * ||x|| * cmax + a_cc * cmax^2,  V_j = ||c_j||^2 - 2 x.c_j
//   |D_j - dist_j| <= g * dist_j,  g = gamma_{d+2}  (D: exact-form distance in F)
// If D_j is minimal then dist_j <= dist_i (1+g)/(1-g) for the approximate argmin
// i, so v_j <= v_i + 2 da + 2 g' dist_i with g' = g/(1-g) and
// dist_i <= v_i + da + ||x||^2.  band = 2 da + 2 g' (v_i + da + ||x||^2).
struct band_consts {
    double a_xc;
    double a_cc;
    double g_canon;   // g' = g / (1 - g)
};

// gamma_m = m*u / (1 - m*u)
__host__ __device__ inline double gamma_m(double m, double u) { return m * u / (1.0 - m * u); }

// Exact-form distance (the contract's D) -- mul and add rounded separately.
template <typename F>
__device__ __forceinline__ F canon_dist(const F* __restrict__ x, const F* __restrict__ c, int d) {
    F acc = F(0);
    for (int f = 0; f < d; ++f) {
        const F diff = x[f] - c[f];
        const F sq = diff * diff;
        acc = acc + sq;
    }
    return acc;
}

// FP64 distance of the objective.
template <typename F>
__device__ __forceinline__ double dist64(const F* __restrict__ x, const F* __restrict__ c, int d) {
    double acc = 0.0;
    for (int f = 0; f < d; ++f) {
        const double diff = double(x[f]) - double(c[f]);
        const double sq = diff * diff;
        acc = acc + sq;
    }
    return acc;
}

// Threshold T for a row: every centroid that can be the exact-form argmin has
// approximate value <= T.  X = ||x|| (rounded up), C = max ||c_j|| (rounded up),
// m1 = smallest approximate value of the row.  The relative and absolute slack
// terms cover the rounding of this computation in double.
__device__ __forceinline__ double band_threshold(const band_consts& b, double X, double C,
                                                 double m1) {
    const double da = b.a_xc * X * C + b.a_cc * C * C;
    const double dist_up = fmax(0.0, m1 + da + X * X);
    const double band = (2.0 * da + 2.0 * b.g_canon * dist_up) * (1.0 + 9.5367431640625e-07);
    return m1 + band + fabs(m1) * 4.44e-16 + (X * X + C * C) * 4.44e-16;
}

}  // namespace km
