// km_refine.hpp -- exact re-score of rows the assignment kernels could not settle.
//
// The reference settles every row from one pass of F-precision GEMM distances
// (kernels_fp_impl.hpp:127-214); this engine settles near-ties in the exact form
// D = sum_f (x_f - c_f)^2 (km_common.hpp) so labels do not depend on which
// matrix-core precision produced the candidates.  Ties go to the lowest index,
// as in select_min_distance (kernels_fp_impl.hpp:198-202).
#pragma once

#include "km_common.hpp"

namespace km {

// One thread per listed row; 2..MAXC candidate ids.
template <typename F>
__global__ __launch_bounds__(256) void refine_candidates_kernel(
    const F* __restrict__ X, int d, const F* __restrict__ C, const refine_lists L,
    int32_t* __restrict__ labels) {
    const int total = L.counters[0];
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const int64_t row = L.cand_rows[e];
        const int cnt = L.cand_count[e];
        const F* x = X + row * d;
        F best = fp_traits<F>::inf;
        int bi = 0x7fffffff;
        for (int c = 0; c < cnt; ++c) {
            const int j = L.cand_ids[int64_t(e) * MAXC + c];
            const F v = canon_dist<F>(x, C + int64_t(j) * d, d);
            if (v < best || (v == best && j < bi)) {
                best = v;
                bi = j;
            }
        }
        labels[row] = bi;
    }
}

// One wave per listed row; lanes stride over all k centroids.  Reproduces the
// sequential scan "best = D_0; for j: if D_j < best" including its behaviour on
// a NaN D_0.
template <typename F>
__global__ __launch_bounds__(256) void refine_full_kernel(const F* __restrict__ X, int d,
                                                          const F* __restrict__ C, int k,
                                                          const refine_lists L,
                                                          int32_t* __restrict__ labels) {
    const int total = L.counters[1];
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    for (int e = wave; e < total; e += nwaves) {
        const int64_t row = L.full_rows[e];
        const F* x = X + row * d;
        F best = fp_traits<F>::inf;
        int bi = 0x7fffffff;
        bool nan0 = false;
        for (int j = lane; j < k; j += 64) {
            const F v = canon_dist<F>(x, C + int64_t(j) * d, d);
            if (j == 0 && v != v) nan0 = true;
            if (v < best) {
                best = v;
                bi = j;
            }
        }
        // wave argmin (value, then index)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const F ob = __shfl_xor(best, off);
            const int oi = __shfl_xor(bi, off);
            if (ob < best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
            }
        }
        const bool anynan0 = __shfl(nan0 ? 1 : 0, 0) != 0;
        if (lane == 0) labels[row] = (anynan0 || bi == 0x7fffffff) ? 0 : bi;
    }
}

}  // namespace km
