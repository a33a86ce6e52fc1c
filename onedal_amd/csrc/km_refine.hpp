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

// Exact re-score of listed rows against their candidate ids (certified
// candidates from the assignment kernels, or ids gathered by the BF16 collect
// pass).  Entry e: row rows[e], counts[e] ids at ids[e * stride].  An entry with
// no ids or more than `stride` (a collect overflow) goes to the full list.
// Each wave takes rb entries: their rows are staged in LDS by all 64 lanes with
// 16-byte loads (row stride d + 16 bytes, so lanes reading different rows hit
// different banks), the (row, id) pairs are spread over the lanes, each lane
// forms the exact distance of its pair (features in order), reading the centroid
// one 32-feature chunk ahead of the sum, and lane r keeps the minimum of entry r
// (ties to the lowest id).
template <typename F>
__host__ __device__ inline int refine_row_stride(int d) {
    return d + int(16 / sizeof(F));
}

template <typename F>
__global__ __launch_bounds__(256) void refine_pairs_kernel(
    const F* __restrict__ X, int d, const F* __restrict__ C, const int32_t* __restrict__ rows,
    const int32_t* __restrict__ counts, const int32_t* __restrict__ ids, int stride,
    const int32_t* __restrict__ count_p, int rb, int32_t* __restrict__ full_rows,
    int32_t* __restrict__ full_counter, int32_t* __restrict__ labels) {
    typedef F F4 __attribute__((ext_vector_type(4)));
    constexpr int CH = 32 / int(sizeof(F));  // 4-vectors per prefetched centroid chunk
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ F pd[4][64];
    __shared__ int pj[4][64];
    __shared__ int pstart[4][64];
    __shared__ int prow[4][64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int ds = refine_row_stride<F>(d);
    F* xs = reinterpret_cast<F*>(smem_raw) + size_t(wave) * rb * ds;
    const int total = *count_p;
    const int nbatch = (total + rb - 1) / rb;
    const int gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    const bool vec = (d % 4) == 0 && (reinterpret_cast<uintptr_t>(C) % (4 * sizeof(F))) == 0 &&
                     (reinterpret_cast<uintptr_t>(X) % (4 * sizeof(F))) == 0;
    const int nv = d / 4;
    for (int b = gw; b < nbatch; b += nw) {
        const int e0 = b * rb;
        const int nr = min(rb, total - e0);
        int my_row = 0, my_cnt = 0;
        if (lane < nr) {
            my_row = rows[e0 + lane];
            my_cnt = counts[e0 + lane];
            if (my_cnt > stride || my_cnt <= 0) {
                full_rows[atomicAdd(full_counter, 1)] = my_row;
                my_cnt = 0;
            }
        }
        prow[wave][lane] = my_row;
        __builtin_amdgcn_wave_barrier();
        // stage the rows: every lane moves 4-vectors (or elements) of any row
        if (vec) {
            for (int i = lane; i < nr * nv; i += 64) {
                const int r = i / nv, q = i - r * nv;
                const F4 v = *reinterpret_cast<const F4*>(X + int64_t(prow[wave][r]) * d + 4 * q);
                *reinterpret_cast<F4*>(xs + r * ds + 4 * q) = v;
            }
        } else {
            for (int i = lane; i < nr * d; i += 64) {
                const int r = i / d, f = i - r * d;
                xs[r * ds + f] = X[int64_t(prow[wave][r]) * d + f];
            }
        }
        // exclusive scan of the pair counts over lanes 0..nr-1
        int incl = my_cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        const int my_start = incl - my_cnt;
        const int npairs = __shfl(incl, 63);
        pstart[wave][lane] = my_start;
        __builtin_amdgcn_wave_barrier();
        F best = fp_traits<F>::inf;
        int bj = 0x7fffffff;
        for (int p0 = 0; p0 < npairs; p0 += 64) {
            const int p = p0 + lane;
            F dist = fp_traits<F>::inf;
            int j = 0x7fffffff;
            if (p < npairs) {
                // entry owning pair p: last r with start[r] <= p
                int r = 0;
                for (int q = 1; q < nr; ++q)
                    if (pstart[wave][q] <= p) r = q;
                j = ids[int64_t(e0 + r) * stride + (p - pstart[wave][r])];
                const F* x = xs + r * ds;
                const F* c = C + int64_t(j) * d;
                F acc = F(0);
                if (vec) {
                    F4 cur[CH], nxt[CH];
#pragma unroll
                    for (int q = 0; q < CH; ++q)
                        cur[q] = q < nv ? *reinterpret_cast<const F4*>(c + 4 * q) : F4{};
                    for (int v0 = 0; v0 < nv; v0 += CH) {
#pragma unroll
                        for (int q = 0; q < CH; ++q)
                            nxt[q] = v0 + CH + q < nv ? *reinterpret_cast<const F4*>(c + 4 * (v0 + CH + q))
                                                      : F4{};
#pragma unroll
                        for (int q = 0; q < CH; ++q) {
                            if (v0 + q < nv) {
                                const F4 x4 = *reinterpret_cast<const F4*>(x + 4 * (v0 + q));
#pragma unroll
                                for (int t = 0; t < 4; ++t) {
                                    const F diff = x4[t] - cur[q][t];
                                    const F sq = diff * diff;
                                    acc = acc + sq;
                                }
                            }
                        }
#pragma unroll
                        for (int q = 0; q < CH; ++q) cur[q] = nxt[q];
                    }
                } else {
                    for (int f = 0; f < d; ++f) {
                        const F diff = x[f] - c[f];
                        const F sq = diff * diff;
                        acc = acc + sq;
                    }
                }
                dist = acc;
            }
            pd[wave][lane] = dist;
            pj[wave][lane] = j;
            __builtin_amdgcn_wave_barrier();
            if (lane < nr && my_cnt > 0) {
                const int q0 = max(my_start, p0), q1 = min(my_start + my_cnt, p0 + 64);
                for (int q = q0; q < q1; ++q) {
                    const F v = pd[wave][q - p0];
                    const int jj = pj[wave][q - p0];
                    if (v < best || (v == best && jj < bj)) {
                        best = v;
                        bj = jj;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (lane < nr && my_cnt > 0) labels[my_row] = bj;
        __builtin_amdgcn_wave_barrier();
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

// refine_full_kernel for d % 4 == 0 with 4-vector aligned X and C.  A workgroup
// takes 32 listed rows (8 per wave), staged in LDS as 4-vectors (feature chunks of
// 32 KB when the rows do not fit); lane l carries centroid j0 + l of a 64-centroid
// tile with its features in registers, a register chunk of 32 bytes per feature
// group loaded one chunk ahead of use, and forms the exact-form sums of the
// wave's 8 rows in feature order, reading the rows as LDS broadcasts.  Same scan
// semantics as refine_full_kernel.
template <typename F>
__global__ __launch_bounds__(256) void refine_full_vec_kernel(const F* __restrict__ X, int d,
                                                              const F* __restrict__ C, int k,
                                                              const refine_lists L,
                                                              int32_t* __restrict__ labels) {
    typedef F F4 __attribute__((ext_vector_type(4)));
    constexpr int RW = 8, RB = 4 * RW;
    constexpr int VT = 32768 / (RB * int(sizeof(F4)));  // 4-vectors per staged row chunk
    constexpr int CQ = 32 / int(sizeof(F));             // 4-vectors per register chunk
    __shared__ F4 Xs[RB][VT];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int total = L.counters[1];
    const int nbatch = (total + RB - 1) / RB;
    const int nv = d / 4;
    const bool one_chunk = nv <= VT;
    const F4* X4 = reinterpret_cast<const F4*>(X);
    auto stage = [&](int b, int v0, int vt) {
        for (int e = tid; e < RB * vt; e += 256) {
            const int r = e / vt, q = e - r * vt;
            const int le = b * RB + r;
            const int64_t row = le < total ? int64_t(L.full_rows[le]) : 0;
            Xs[r][q] = X4[row * nv + v0 + q];
        }
    };
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
        if (one_chunk) {
            __syncthreads();
            stage(b, 0, nv);
            __syncthreads();
        }
        for (int j0 = 0; j0 < k; j0 += 64) {
            const int j = j0 + lane;
            const F4* c = reinterpret_cast<const F4*>(C + int64_t(j < k ? j : 0) * d);
            F acc[RW];
#pragma unroll
            for (int r = 0; r < RW; ++r) acc[r] = F(0);
            for (int v0 = 0; v0 < nv; v0 += VT) {
                const int vt = min(VT, nv - v0);
                if (!one_chunk) {
                    __syncthreads();
                    stage(b, v0, vt);
                    __syncthreads();
                }
                F4 cur[CQ], nxt[CQ];
#pragma unroll
                for (int q = 0; q < CQ; ++q) cur[q] = q < vt ? c[v0 + q] : F4{};
                for (int u0 = 0; u0 < vt; u0 += CQ) {
#pragma unroll
                    for (int q = 0; q < CQ; ++q)
                        nxt[q] = u0 + CQ + q < vt ? c[v0 + u0 + CQ + q] : F4{};
#pragma unroll
                    for (int q = 0; q < CQ; ++q) {
                        if (u0 + q < vt) {
#pragma unroll
                            for (int r = 0; r < RW; ++r) {
                                const F4 x4 = Xs[wave * RW + r][u0 + q];
#pragma unroll
                                for (int t = 0; t < 4; ++t) {
                                    const F diff = x4[t] - cur[q][t];
                                    const F sq = diff * diff;
                                    acc[r] = acc[r] + sq;
                                }
                            }
                        }
                    }
#pragma unroll
                    for (int q = 0; q < CQ; ++q) cur[q] = nxt[q];
                }
            }
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
