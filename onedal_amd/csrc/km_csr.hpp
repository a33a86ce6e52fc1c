// km_csr.hpp -- kernels of the CSR Lloyd path (method lloyd_csr).
//
// Replaces, on the oneDAL side (backend/gpu/kernels_csr_impl.hpp):
//   compute_data_squares            :38-55    -> csr_row_squares_kernel
//   transpose + compute_squares     :70-92, kernels_fp_impl.hpp:402-414
//                                   -> csr_transpose_kernel, csr_centroid_squares_kernel
//   assign_clusters (sparse gemm, select, complete_closest_distances)
//                                   :118-187  -> csr_assign_kernel
//   calc_objective_function         :193-203  -> csr_objective_chunks_kernel
//   update_centroids (local atomics):219-322  -> csr_chunk_sums_kernel + finalize_centroids_kernel
//   copy_candidates_from_data       empty_cluster_handling_dpc.cpp:152-189 -> csr_fill_empty_kernel
//
// Arithmetic (fixed order, so results are identical run to run and equal to the
// oracle's restatement; the reference's sparse gemm and atomics fix no order):
//   xs_i  = sum over the row's stored entries, in stored order, of v*v      (F)
//   cs_j  = sum_f c_jf*c_jf, features in order                              (F)
//   dot   = sum over stored entries, in stored order, of v*c_j[col]         (F, no FMA)
//   s_ij  = cs_j - 2*dot (F); label = first j with the smallest s (a NaN s_i0
//           keeps label 0, other NaNs never win); closest_i = s_i,label + xs_i (F)
//   sums  = the dense path's chunk sums (REDUCE_CHUNK members, reduce_groups
//           partials, double), adding stored entries in member order; an absent
//           entry would add +0.0, so on the same rows the sums equal the dense ones
//   objective = sum of double(closest_i) in the dense objective's order.
// Column indices need not be sorted; duplicate entries of a row are added one by
// one in stored order (sums) or overwrite in stored order (empty-cluster fill).
#pragma once

#include "km_common.hpp"

namespace km {

// Structure check: row offsets non-decreasing from `base` to base + nnz, every
// column index in [base, base + d).  Each thread reads a row's entries only after
// its own offsets passed the check.  flag bit 1: bad offsets, bit 2: bad column,
// bit 4: a row whose column indices do not increase strictly (unsorted or
// repeated; allowed, reported for the initialisation's feature-order walk).
__global__ __launch_bounds__(256) void csr_check_kernel(const int64_t* __restrict__ cols,
                                                        const int64_t* __restrict__ offs,
                                                        int64_t base, int64_t n, int64_t d,
                                                        int64_t nnz, int32_t* __restrict__ flag) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t o0 = offs[i] - base, o1 = offs[i + 1] - base;
        bool ok = o0 >= 0 && o1 >= o0 && o1 <= nnz;
        if (i == 0 && o0 != 0) ok = false;
        if (i == n - 1 && o1 != nnz) ok = false;
        if (!ok) {
            atomicOr(flag, 1);
            continue;
        }
        int64_t prev = -1;
        bool inc = true;
        for (int64_t p = o0; p < o1; ++p) {
            const int64_t c = cols[p] - base;
            if (c < 0 || c >= d) {
                atomicOr(flag, 2);
                break;
            }
            inc = inc && c > prev;
            prev = c;
        }
        if (!inc) atomicOr(flag, 4);
    }
}

// xs_i: squares of a row's stored entries, summed in stored order
template <typename F>
__global__ __launch_bounds__(256) void csr_row_squares_kernel(const F* __restrict__ vals,
                                                              const int64_t* __restrict__ offs,
                                                              int64_t base, int64_t n,
                                                              F* __restrict__ xs) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t o0 = offs[i] - base, o1 = offs[i + 1] - base;
        F acc(0);
        for (int64_t p = o0; p < o1; ++p) {
            const F v = vals[p];
            const F sq = v * v;
            acc = acc + sq;
        }
        xs[i] = acc;
    }
}

// Ct (d x kp) = C^T, rows padded with zeros to kp (a multiple of the 16-byte
// vector), for coalesced vector reads over centroids: 64 x 64 tiles through LDS
// (grid: ceil(d/64) x up to 65535 workgroups of 256 threads, striding over k)
template <typename F>
__global__ __launch_bounds__(256) void csr_transpose_kernel(const F* __restrict__ C, int64_t k,
                                                            int64_t kp, int64_t d, F* __restrict__ Ct) {
    __shared__ F tile[64][65];
    const int64_t f0 = int64_t(blockIdx.x) * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int64_t j0 = int64_t(blockIdx.y) * 64; j0 < kp; j0 += int64_t(gridDim.y) * 64) {
        for (int r = ty; r < 64; r += 4) {
            const int64_t j = j0 + r, f = f0 + tx;
            if (j < k && f < d) tile[r][tx] = C[j * d + f];
        }
        __syncthreads();
        for (int r = ty; r < 64; r += 4) {
            const int64_t f = f0 + r, j = j0 + tx;
            if (j < kp && f < d) Ct[f * kp + j] = j < k ? tile[tx][r] : F(0);
        }
        __syncthreads();
    }
}

// cs_j = |c_j|^2, features in order: one wave per centroid copies the row to LDS
// (coalesced), then lane 0 adds the squares in feature order
template <typename F>
__global__ __launch_bounds__(64) void csr_centroid_squares_kernel(const F* __restrict__ C, int64_t k,
                                                                  int d, F* __restrict__ cs) {
    extern __shared__ unsigned char sq_raw[];
    F* row = reinterpret_cast<F*>(sq_raw);  // [d]
    for (int64_t j = blockIdx.x; j < k; j += gridDim.x) {
        for (int f = threadIdx.x; f < d; f += 64) row[f] = C[j * d + f];
        __syncthreads();
        if (threadIdx.x == 0) {
            F acc(0);
            for (int f = 0; f < d; ++f) {
                const F c = row[f];
                const F sq = c * c;
                acc = acc + sq;
            }
            cs[j] = acc;
        }
        __syncthreads();
    }
}

// (s, j) candidate order: the smaller s wins, equal values go to the lower j;
// NaN never wins (has == false stands for "no candidate yet")
template <typename F>
__device__ __forceinline__ void csr_take(F& bs, int& bj, bool& has, F s, int j) {
    if (s != s) return;
    if (!has || s < bs || (s == bs && j < bj)) {
        bs = s;
        bj = j;
        has = true;
    }
}

// value of lane u (wave-uniform u) in scalar registers
__device__ __forceinline__ float lane_value(float x, int u) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), u));
}
__device__ __forceinline__ double lane_value(double x, int u) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane(int(b & 0xFFFFFFFFll), u);
    const int hi = __builtin_amdgcn_readlane(int(b >> 32), u);
    return __longlong_as_double((long long)(unsigned(lo)) | ((long long)hi << 32));
}

// 16-byte vectors of F
template <typename F>
struct csr_vec;
template <>
struct csr_vec<float> {
    static constexpr int N = 4;
    typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct csr_vec<double> {
    static constexpr int N = 2;
    typedef double type __attribute__((ext_vector_type(2)));
};

// One wave per row.  Lane l accumulates the dots of the VEC consecutive
// centroids jb + VEC (l + 64 t) + e (t < JT, e < VEC) over the row's entries,
// which are loaded 64 at a time and taken in stored order from their lanes;
// each entry reads one 16-byte vector of its transposed-centroid row per t.
// Then each lane's best and a wave reduction (ties to the lower j).
template <typename F, int JT>
__global__ __launch_bounds__(256) void csr_assign_kernel(
    const F* __restrict__ vals, const int64_t* __restrict__ cols, const int64_t* __restrict__ offs,
    int64_t base, int64_t n, int k, int kp, const F* __restrict__ Ct, const F* __restrict__ cs,
    const F* __restrict__ xs, int32_t* __restrict__ labels, F* __restrict__ closest) {
    constexpr int VEC = csr_vec<F>::N;
    typedef typename csr_vec<F>::type V;
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = (int64_t(gridDim.x) * blockDim.x) >> 6;
    for (int64_t i = wave; i < n; i += nwaves) {
        const int64_t o0 = offs[i] - base, o1 = offs[i + 1] - base;
        F bs(0);
        int bj = 0;
        bool has = false;
        F s0(0);  // s of centroid 0 (lane 0, first block)
        for (int jb = 0; jb < k; jb += 64 * VEC * JT) {
            F acc[JT][VEC];
#pragma unroll
            for (int t = 0; t < JT; ++t)
#pragma unroll
                for (int e = 0; e < VEC; ++e) acc[t][e] = F(0);
            for (int64_t p0 = o0; p0 < o1; p0 += 64) {
                const int cnt = int(min<int64_t>(64, o1 - p0));
                int my_c = 0;  // a column index (< d <= 4096)
                F my_v(0);
                if (lane < cnt) {
                    my_c = int(cols[p0 + lane] - base);
                    my_v = vals[p0 + lane];
                }
                for (int u = 0; u < cnt; ++u) {
                    // entry u of the group, read from its lane into scalar registers
                    const int c = __builtin_amdgcn_readlane(my_c, u);
                    const F v = lane_value(my_v, u);
                    const F* crow = Ct + int64_t(c) * kp;
#pragma unroll
                    for (int t = 0; t < JT; ++t) {
                        const int j = jb + VEC * (64 * t + lane);
                        if (j < k) {
                            const V cv = *reinterpret_cast<const V*>(crow + j);
#pragma unroll
                            for (int e = 0; e < VEC; ++e) {
                                const F pr = v * cv[e];
                                acc[t][e] = acc[t][e] + pr;
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < JT; ++t)
#pragma unroll
                for (int e = 0; e < VEC; ++e) {
                    const int j = jb + VEC * (64 * t + lane) + e;
                    if (j < k) {
                        const F two_dot = F(2) * acc[t][e];
                        const F s = cs[j] - two_dot;
                        if (j == 0) s0 = s;
                        csr_take(bs, bj, has, s, j);
                    }
                }
        }
        // wave reduction of (bs, bj, has)
        for (int off = 32; off > 0; off >>= 1) {
            const F os = __shfl_xor(bs, off);
            const int oj = __shfl_xor(bj, off);
            const int oh = __shfl_xor(int(has), off);
            if (oh) csr_take(bs, bj, has, os, oj);
        }
        s0 = __shfl(s0, 0);
        if (lane == 0) {
            int lab = bj;
            F best = bs;
            if (s0 != s0 || !has) {  // s_i0 NaN, or every s NaN: label 0
                lab = 0;
                best = s0;
            }
            labels[i] = lab;
            closest[i] = best + xs[i];
        }
    }
}

// objective over closest_i (as double) in the dense objective's order; also
// writes the per-row keys used to pick empty-cluster candidates (and subtracted
// from the objective for each one taken)
template <typename F>
__global__ __launch_bounds__(256) void csr_objective_chunks_kernel(const F* __restrict__ closest,
                                                                   int64_t n,
                                                                   double* __restrict__ d64_out,
                                                                   double* __restrict__ chunk_sums,
                                                                   const int32_t* __restrict__ gate) {
    if (gate && gate[0] == 0) return;  // gated off (no empty clusters this iteration)
    __shared__ double part[256];
    const int t = threadIdx.x;
    const int64_t nchunks = (n + OBJ_CHUNK - 1) / OBJ_CHUNK;
    for (int64_t q = blockIdx.x; q < nchunks; q += gridDim.x) {
        const int64_t r0 = q * OBJ_CHUNK;
        const int rows = int(min<int64_t>(OBJ_CHUNK, n - r0));
        double p = 0.0;
        for (int r = t; r < rows; r += 256) {
            const double v = double(closest[r0 + r]);
            // candidate keys: a rounding-negative distance counts as 0 (the
            // radix select orders the raw bits, which is numeric only for +0 and up)
            if (d64_out) d64_out[r0 + r] = (v != v || v > 0.0) ? v : 0.0;
            p = p + v;
        }
        part[t] = p;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (t < s) part[t] = part[t] + part[t + s];
            __syncthreads();
        }
        if (t == 0) chunk_sums[q] = part[0];
        __syncthreads();
    }
}

// members whose entries are loaded together by the chunk-sum kernel
constexpr int CSR_AHEAD = 8;

__device__ __forceinline__ int64_t lane_value(int64_t x, int u) {
    const int lo = __builtin_amdgcn_readlane(int(x & 0xFFFFFFFFll), u);
    const int hi = __builtin_amdgcn_readlane(int(x >> 32), u);
    return int64_t((unsigned long long)(unsigned(lo)) | ((unsigned long long)(unsigned(hi)) << 32));
}

// Chunk sums of the CSR rows: one single-wave workgroup per chunk of
// REDUCE_CHUNK members, G = reduce_groups(d) partial vectors of d doubles in LDS
// (member m of the chunk goes to partial m % G).  A row whose column indices are
// strictly increasing is added by all lanes at once (distinct columns); any
// other row by lane 0 alone, entry by entry, so every column is summed in member
// order and, inside a row, in stored order.  Partials are added in g order.
// The wave's LDS operations execute in issue order, so a member's adds see the
// previous member's.  Groups of 64 members whose rows all hold at most 64
// entries load CSR_AHEAD rows' entries at a time from registers (the offsets of
// the whole group first); other groups go member by member.
template <typename F>
__global__ __launch_bounds__(64) void csr_chunk_sums_kernel(
    const F* __restrict__ vals, const int64_t* __restrict__ cols, const int64_t* __restrict__ offs,
    int64_t base, int d, int k, const int32_t* __restrict__ perm,
    const int64_t* __restrict__ offsets, const int32_t* __restrict__ chunk_base,
    const int32_t* __restrict__ meta, double* __restrict__ part) {
    extern __shared__ double acc[];  // [G * d]
    const int lane = threadIdx.x;
    const int total = meta[1];
    const int G = reduce_groups(d);
    const int width = G * d;
    for (int q = blockIdx.x; q < total; q += gridDim.x) {
        int lo = 0, hi = k;  // chunk_base[k] = total > q
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (chunk_base[mid] <= q) lo = mid; else hi = mid;
        }
        const int j = lo;
        const int64_t m_begin = offsets[j] + int64_t(q - chunk_base[j]) * REDUCE_CHUNK;
        const int64_t m_end = min(m_begin + REDUCE_CHUNK, offsets[j + 1]);
        for (int e = lane; e < width; e += 64) acc[e] = 0.0;
        __syncthreads();
        // groups of up to 64 members: lane t holds member t's row and offsets
        for (int64_t m0 = m_begin; m0 < m_end; m0 += 64) {
            const int cnt = int(min<int64_t>(64, m_end - m0));
            int64_t my_o0 = 0;
            int my_len = 0;
            if (lane < cnt) {
                const int64_t row = perm[m0 + lane];
                my_o0 = offs[row] - base;
                my_len = int(offs[row + 1] - base - my_o0);
            }
            if (__all(my_len <= 64)) {
                // every row fits the lanes: the entries of CSR_AHEAD members are
                // loaded together, then added member by member
                for (int u0 = 0; u0 < cnt; u0 += CSR_AHEAD) {
                    int cc[CSR_AHEAD], ln[CSR_AHEAD];
                    F vv[CSR_AHEAD];
#pragma unroll
                    for (int q = 0; q < CSR_AHEAD; ++q) {
                        cc[q] = 0;
                        vv[q] = F(0);
                        ln[q] = 0;
                        if (u0 + q < cnt) {
                            ln[q] = __builtin_amdgcn_readlane(my_len, u0 + q);
                            const int64_t o0 = lane_value(my_o0, u0 + q);
                            if (lane < ln[q]) {
                                cc[q] = int(cols[o0 + lane] - base);
                                vv[q] = vals[o0 + lane];
                            }
                        }
                    }
#pragma unroll
                    for (int q = 0; q < CSR_AHEAD; ++q) {
                        if (u0 + q < cnt) {
                            const int len = ln[q];
                            double* pg = acc + int((m0 + u0 + q - m_begin) % G) * d;
                            const int prev = __shfl_up(cc[q], 1);
                            const bool inc = lane == 0 || lane >= len || prev < cc[q];
                            if (__all(inc)) {
                                // strictly increasing columns: distinct, added at once
                                if (lane < len) pg[cc[q]] = pg[cc[q]] + double(vv[q]);
                            } else {
                                for (int e = 0; e < len; ++e) {
                                    const int c = __builtin_amdgcn_readlane(cc[q], e);
                                    const F v = lane_value(vv[q], e);
                                    if (lane == 0) pg[c] = pg[c] + double(v);
                                }
                            }
                        }
                    }
                }
                continue;
            }
            for (int64_t m = m0; m < m0 + cnt; ++m) {
                const int64_t row = perm[m];
                double* pg = acc + int((m - m_begin) % G) * d;
                const int64_t o0 = offs[row] - base, o1 = offs[row + 1] - base;
                bool sorted = true;
                for (int64_t p = o0 + 1 + lane; p < o1; p += 64)
                    if (cols[p] <= cols[p - 1]) sorted = false;
                sorted = __all(sorted);
                if (sorted) {
                    for (int64_t p = o0 + lane; p < o1; p += 64) {
                        const int64_t c = cols[p] - base;
                        pg[c] = pg[c] + double(vals[p]);
                    }
                } else if (lane == 0) {
                    for (int64_t p = o0; p < o1; ++p) {
                        const int64_t c = cols[p] - base;
                        pg[c] = pg[c] + double(vals[p]);
                    }
                }
                __syncthreads();
            }
        }
        for (int f = lane; f < d; f += 64) {
            double t = acc[f];
            for (int g = 1; g < G; ++g) t = t + acc[g * d + f];
            part[int64_t(q) * d + f] = t;
        }
        __syncthreads();
    }
}

// empty cluster e takes candidate row cand[e]: zeros, then the stored entries in
// stored order (a later duplicate overwrites)
template <typename F>
__global__ __launch_bounds__(64) void csr_fill_empty_kernel(
    const F* __restrict__ vals, const int64_t* __restrict__ cols, const int64_t* __restrict__ offs,
    int64_t base, int d, const int32_t* __restrict__ empty_list, const int32_t* __restrict__ cand,
    const int32_t* __restrict__ meta, F* __restrict__ Cnext) {
    const int m = meta[0];  // empty clusters this iteration (picks in cand, ec_order_kernel)
    for (int e = blockIdx.x; e < m; e += gridDim.x) {
        F* dst = Cnext + int64_t(empty_list[e]) * d;
        for (int f = threadIdx.x; f < d; f += blockDim.x) dst[f] = F(0);
        __syncthreads();
        if (threadIdx.x == 0) {
            const int64_t row = cand[e];
            const int64_t o0 = offs[row] - base, o1 = offs[row + 1] - base;
            for (int64_t p = o0; p < o1; ++p) dst[cols[p] - base] = vals[p];
        }
        __syncthreads();
    }
}

}  // namespace km
