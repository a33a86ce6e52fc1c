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
    for (int64_t r0 = int64_t(blockIdx.x) * 256; r0 < n; r0 += int64_t(gridDim.x) * 256) {
        const int nr = int(min<int64_t>(256, n - r0));
        double acc = 0.0;
        for (int f0 = 0; f0 < d; f0 += INIT_FT) {
            const int ft = min(INIT_FT, d - f0);
            __syncthreads();
            for (int e = t; e < nr * ft; e += 256) {
                const int r = e / ft, q = e - r * ft;
                Xs[r][q] = X[(r0 + r) * d + f0 + q];
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

// the sampled row (one thread): see the file comment
__global__ void init_select_kernel(const double* __restrict__ D2, int64_t n,
                                   const double* __restrict__ chunk_sums, double u,
                                   int64_t* __restrict__ chosen) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int64_t nchunks = (n + OBJ_CHUNK - 1) / OBJ_CHUNK;
    double total = 0.0;
    for (int64_t q = 0; q < nchunks; ++q) total = total + chunk_sums[q];
    int64_t idx;
    if (!(total > 0.0)) {
        const double pos = u * double(n);
        idx = pos >= double(n - 1) ? n - 1 : int64_t(pos);
    } else {
        const double target = u * total;
        double cum = 0.0;
        int64_t q = 0;
        for (; q + 1 < nchunks; ++q) {
            if (cum + chunk_sums[q] > target) break;
            cum = cum + chunk_sums[q];
        }
        const int64_t r0 = q * OBJ_CHUNK;
        const int64_t r1 = min<int64_t>(n, r0 + OBJ_CHUNK);
        idx = -1;
        int64_t last_pos = -1;
        double s = cum;
        for (int64_t r = r0; r < r1; ++r) {
            const double w = D2[r];
            if (w > 0.0) last_pos = r;
            s = s + w;
            if (w > 0.0 && s > target) {
                idx = r;
                break;
            }
        }
        if (idx < 0) idx = last_pos >= 0 ? last_pos : r0;
    }
    chosen[0] = idx;
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
