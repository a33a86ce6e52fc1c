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

// Rows re-scored against every centroid.  A workgroup takes 16 listed rows
// (4 per wave); centroid tiles of 64 (one per lane) x FT features are staged
// transposed in LDS, the rows' features beside them, and each lane carries the
// exact-form sum of its centroid for the wave's 4 rows across feature tiles (in
// feature order).  Reproduces the sequential scan "best = D_0; for j: if D_j <
// best", including its behaviour on a NaN D_0.
template <typename F>
__global__ __launch_bounds__(256) void refine_full_kernel(const F* __restrict__ X, int d,
                                                          const F* __restrict__ C, int k,
                                                          const refine_lists L,
                                                          int32_t* __restrict__ labels) {
    constexpr int KT = 64, RW = 4, RB = 4 * RW;
    constexpr int FT = sizeof(F) == 4 ? 112 : 56;
    __shared__ F Ct[FT][KT + 1];
    __shared__ F Xs[RB][FT];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int total = L.counters[1];
    const int nbatch = (total + RB - 1) / RB;
    for (int b = blockIdx.x; b < nbatch; b += gridDim.x) {
        int64_t rows[RW];
        F best[RW];
        int bi[RW];
        bool nan0[RW];
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const int e = b * RB + wave * RW + r;
            rows[r] = e < total ? int64_t(L.full_rows[e]) : -1;
            best[r] = fp_traits<F>::inf;
            bi[r] = 0x7fffffff;
            nan0[r] = false;
        }
        for (int j0 = 0; j0 < k; j0 += KT) {
            F acc[RW];
#pragma unroll
            for (int r = 0; r < RW; ++r) acc[r] = F(0);
            for (int f0 = 0; f0 < d; f0 += FT) {
                const int ft = min(FT, d - f0);
                __syncthreads();
                for (int e = tid; e < KT * ft; e += 256) {
                    const int j = e / ft, f = e - (e / ft) * ft;
                    Ct[f][j] = (j0 + j < k) ? C[int64_t(j0 + j) * d + f0 + f] : F(0);
                }
                for (int e = tid; e < RB * ft; e += 256) {
                    const int r = e / ft, f = e - (e / ft) * ft;
                    const int le = b * RB + r;
                    const int64_t row = le < total ? int64_t(L.full_rows[le]) : 0;
                    Xs[r][f] = X[row * d + f0 + f];
                }
                __syncthreads();
                for (int f = 0; f < ft; ++f) {
                    const F c = Ct[f][lane];
#pragma unroll
                    for (int r = 0; r < RW; ++r) {
                        const F diff = Xs[wave * RW + r][f] - c;
                        const F sq = diff * diff;
                        acc[r] = acc[r] + sq;
                    }
                }
            }
            const int j = j0 + lane;
            if (j < k) {
#pragma unroll
                for (int r = 0; r < RW; ++r) {
                    if (j == 0 && acc[r] != acc[r]) nan0[r] = true;
                    if (acc[r] < best[r]) {
                        best[r] = acc[r];
                        bi[r] = j;
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            F bv = best[r];
            int bj = bi[r];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const F ob = __shfl_xor(bv, off);
                const int oi = __shfl_xor(bj, off);
                if (ob < bv || (ob == bv && oi < bj)) {
                    bv = ob;
                    bj = oi;
                }
            }
            const bool n0 = __shfl(nan0[r] ? 1 : 0, 0) != 0;
            if (lane == 0 && rows[r] >= 0) labels[rows[r]] = (n0 || bj == 0x7fffffff) ? 0 : bj;
        }
    }
}

}  // namespace km
