// km_init.hpp -- initial centroids on the device (kmeans_init).
//
// Replaces kmeans_init's GPU kernels (cpp/oneapi/dal/algo/kmeans_init/backend/gpu/
// compute_kernels_impl.hpp:36-68 for dense, compute_kernel_distr_random_dense_dpc.cpp
// for random_dense) and adds plus_plus_dense, which the reference runs only on the
// host (its GPU kernel throws unimplemented; kmeans train without initial
// centroids calls DAAL's k-means++ on the CPU, train_kernel_lloyd_dense_dpc.cpp:44-74).
//
// k-means++ (D^2 sampling) here, in a fixed arithmetic order so that the oracle
// restates it bit for bit:
//   D2_r = min over chosen centroids of sum_f ((double)x_rf - (double)c_f)^2
//          (features in order, each sub/mul/add rounded separately);
//   total = chunk sums of OBJ_CHUNK rows (thread t of 256 sums rows t, t+256, ...
//          in order, then a tree of the 256 partials), chunks added in order;
//   target = u * total with u from the counter RNG; the chosen row is the first
//          whose running sum (chunks in order, then that chunk's rows in order)
//          exceeds target; if rounding leaves none, the chunk's last row with
//          D2 > 0; if total is 0, row min(n - 1, floor(u * n)).
#pragma once

#include <cstdint>
#include <type_traits>

#include "km_common.hpp"

namespace km {

constexpr int INIT_FT = 32;  // features staged per pass

// D2[r] = dist64(x_r, c) (first) or min(D2[r], dist64(x_r, c)); 256 rows per
// workgroup staged through LDS 32 features at a time, each thread summing its
// own row in feature order.  c points at the chosen centroid (device memory).
template <typename F>
__global__ __launch_bounds__(256) void init_d2_kernel(const F* __restrict__ X, int64_t n, int d,
                                                      const F* __restrict__ c, double* __restrict__ D2,
                                                      int first) {
    __shared__ F Xs[256][INIT_FT + 1];
    const int t = threadIdx.x;
    // float rows with d % 4 == 0 and a 16-byte aligned base load as 4-vectors
    typedef float F4 __attribute__((ext_vector_type(4)));
    const bool vec = std::is_same<F, float>::value && d % 4 == 0 &&
                     reinterpret_cast<uintptr_t>(X) % 16 == 0;
    for (int64_t r0 = int64_t(blockIdx.x) * 256; r0 < n; r0 += int64_t(gridDim.x) * 256) {
        const int nr = int(min<int64_t>(256, n - r0));
        double acc = 0.0;
        for (int f0 = 0; f0 < d; f0 += INIT_FT) {
            const int ft = min(INIT_FT, d - f0);
            __syncthreads();
            if (vec) {
                const int nv = ft / 4;
                for (int e = t; e < nr * nv; e += 256) {
                    const int r = e / nv, q = e - r * nv;
                    const F4 v = *reinterpret_cast<const F4*>(reinterpret_cast<const float*>(X) + (r0 + r) * d +
                                                              f0 + 4 * q);
                    Xs[r][4 * q] = F(v[0]);
                    Xs[r][4 * q + 1] = F(v[1]);
                    Xs[r][4 * q + 2] = F(v[2]);
                    Xs[r][4 * q + 3] = F(v[3]);
                }
            } else {
                for (int e = t; e < nr * ft; e += 256) {
                    const int r = e / ft, q = e - r * ft;
                    Xs[r][q] = X[(r0 + r) * d + f0 + q];
                }
            }
            __syncthreads();
            if (t < nr) {
                for (int q = 0; q < ft; ++q) {
                    const double diff = double(Xs[t][q]) - double(c[f0 + q]);
                    const double sq = diff * diff;
                    acc = acc + sq;
                }
            }
        }
        if (t < nr) {
            const int64_t r = r0 + t;
            D2[r] = first ? acc : fmin(D2[r], acc);
        }
    }
}

// chunk sums of v in the order of the file comment (one workgroup per chunk)
__global__ __launch_bounds__(256) void init_chunk_sums_kernel(const double* __restrict__ v, int64_t n,
                                                              double* __restrict__ chunk_sums) {
    __shared__ double part[256];
    const int t = threadIdx.x;
    const int64_t nchunks = (n + OBJ_CHUNK - 1) / OBJ_CHUNK;
    for (int64_t q = blockIdx.x; q < nchunks; q += gridDim.x) {
        const int64_t r0 = q * OBJ_CHUNK;
        const int rows = int(min<int64_t>(OBJ_CHUNK, n - r0));
        double p = 0.0;
        for (int r = t; r < rows; r += 256) p = p + v[r0 + r];
        __syncthreads();
        part[t] = p;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (t < s) part[t] = part[t] + part[t + s];
            __syncthreads();
        }
        if (t == 0) chunk_sums[q] = part[0];
    }
}

// The sampled row (one workgroup of 256): see the file comment.  The sequential
// walks are done by thread 0 over LDS tiles of 4096 values that the whole
// workgroup loads, so the order of every sum is the sequential one.
__global__ __launch_bounds__(256) void init_select_kernel(const double* __restrict__ D2, int64_t n,
                                                         const double* __restrict__ chunk_sums, double u,
                                                         int64_t* __restrict__ chosen) {
    constexpr int TILE = 4096;
    __shared__ double tile[TILE];
    __shared__ double sh_total, sh_cum;
    __shared__ int64_t sh_q;
    const int t = threadIdx.x;
    const int64_t nchunks = (n + OBJ_CHUNK - 1) / OBJ_CHUNK;
    // total = chunk sums added in order
    double total = 0.0;
    for (int64_t b = 0; b < nchunks; b += TILE) {
        const int m = int(min<int64_t>(TILE, nchunks - b));
        __syncthreads();
        for (int i = t; i < m; i += 256) tile[i] = chunk_sums[b + i];
        __syncthreads();
        if (t == 0) {
            // sixteen values in registers, then their adds in order
            int i = 0;
            for (; i + 16 <= m; i += 16) {
                double v[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) v[j] = tile[i + j];
#pragma unroll
                for (int j = 0; j < 16; ++j) total = total + v[j];
            }
            for (; i < m; ++i) total = total + tile[i];
        }
    }
    if (t == 0) sh_total = total;
    __syncthreads();
    total = sh_total;
    if (!(total > 0.0)) {
        if (t == 0) {
            const double pos = u * double(n);
            chosen[0] = pos >= double(n - 1) ? n - 1 : int64_t(pos);
        }
        return;
    }
    const double target = u * total;
    // chunk q: the first whose running sum exceeds target (the last chunk otherwise)
    double cum = 0.0;
    int64_t qsel = -1;
    for (int64_t b = 0; b < nchunks; b += TILE) {
        const int m = int(min<int64_t>(TILE, nchunks - b));
        __syncthreads();
        for (int i = t; i < m; i += 256) tile[i] = chunk_sums[b + i];
        __syncthreads();
        if (t == 0) {
            for (int i0 = 0; i0 < m && qsel < 0; i0 += 16) {
                double v[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) v[j] = i0 + j < m ? tile[i0 + j] : 0.0;
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const int64_t q = b + i0 + j;
                    if (qsel < 0 && i0 + j < m) {
                        if (q == nchunks - 1 || cum + v[j] > target)
                            qsel = q;
                        else
                            cum = cum + v[j];
                    }
                }
            }
        }
    }
    if (t == 0) {
        sh_q = qsel;
        sh_cum = cum;
    }
    __syncthreads();
    qsel = sh_q;
    cum = sh_cum;
    // the row inside chunk qsel
    const int64_t r0 = qsel * OBJ_CHUNK;
    const int rows = int(min<int64_t>(OBJ_CHUNK, n - r0));
    static_assert(OBJ_CHUNK <= TILE, "one chunk fits the tile");
    for (int i = t; i < rows; i += 256) tile[i] = D2[r0 + i];
    __syncthreads();
    if (t == 0) {
        int64_t idx = -1, last_pos = -1;
        double s = cum;
        for (int i0 = 0; i0 < rows && idx < 0; i0 += 16) {
            double v[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = i0 + j < rows ? tile[i0 + j] : 0.0;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                if (idx < 0 && i0 + j < rows) {
                    const double w = v[j];
                    if (w > 0.0) last_pos = r0 + i0 + j;
                    s = s + w;
                    if (w > 0.0 && s > target) idx = r0 + i0 + j;
                }
            }
        }
        if (idx < 0) idx = last_pos >= 0 ? last_pos : r0;
        chosen[0] = idx;
    }
}

// C[row_out] = X[*chosen] (or X[fixed] when chosen is null)
template <typename F>
__global__ void init_copy_row_kernel(const F* __restrict__ X, int d, const int64_t* __restrict__ chosen,
                                     int64_t fixed, F* __restrict__ C, int64_t row_out) {
    const int64_t src = chosen ? chosen[0] : fixed;
    for (int f = threadIdx.x; f < d; f += blockDim.x) C[row_out * d + f] = X[src * d + f];
}

// C[i] = X[ids[i]] for i < k (random_dense); one workgroup per row
template <typename F>
__global__ void init_gather_rows_kernel(const F* __restrict__ X, int d, const int64_t* __restrict__ ids,
                                        int64_t k, F* __restrict__ C) {
    for (int64_t i = blockIdx.x; i < k; i += gridDim.x)
        for (int f = threadIdx.x; f < d; f += blockDim.x) C[i * d + f] = X[ids[i] * d + f];
}

}  // namespace km
