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
//   chunk sums of OBJ_CHUNK rows (thread t of 256 sums rows t, t+256, ... in
//          order, then a tree of the 256 partials); group sums of INIT_GROUP
//          chunk sums each (in order); total = the group sums in order;
//   target = u * total with u from the counter RNG.  The walk descends with one
//          running sum: the first group whose running sum exceeds target (the
//          last group otherwise), then in it the first such chunk (its last
//          chunk otherwise), then in that chunk the first such block of
//          INIT_SUB rows (block sums in row order; its last block otherwise),
//          then in that block the first row with D2 > 0 whose running sum
//          exceeds target.  Rounding fallbacks: the block's last row with
//          D2 > 0, else the chunk's, else the chunk's first row.  If total is
//          0: row min(n - 1, floor(u * n)).
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

constexpr int INIT_GROUP = 64;  // chunk sums per group sum
constexpr int INIT_SUB = 16;    // rows per block of the last walk level

// The sampled row (one workgroup of 256): see the file comment.  Group sums and
// block sums are formed by all threads, each summing its own range in order;
// thread 0 does the short sequential walks.
__global__ __launch_bounds__(256) void init_select_kernel(const double* __restrict__ D2, int64_t n,
                                                         const double* __restrict__ chunk_sums, double u,
                                                         int64_t* __restrict__ chosen) {
    constexpr int GT = 2048;  // group sums per LDS tile
    static_assert(OBJ_CHUNK / INIT_SUB <= 256, "one block sum per thread");
    __shared__ double gsum[GT];
    __shared__ double bsum[OBJ_CHUNK / INIT_SUB];
    __shared__ double sh_total, sh_cum;
    __shared__ int64_t sh_sel;
    const int t = threadIdx.x;
    const int64_t nchunks = (n + OBJ_CHUNK - 1) / OBJ_CHUNK;
    const int64_t ngroups = (nchunks + INIT_GROUP - 1) / INIT_GROUP;
    auto group_sum = [&](int64_t g) {
        const int64_t q0 = g * INIT_GROUP, q1 = min<int64_t>(nchunks, q0 + INIT_GROUP);
        double sg = 0.0;
        for (int64_t q = q0; q < q1; ++q) sg = sg + chunk_sums[q];
        return sg;
    };
    // total = group sums in order
    double total = 0.0;
    for (int64_t b = 0; b < ngroups; b += GT) {
        const int m = int(min<int64_t>(GT, ngroups - b));
        __syncthreads();
        for (int i = t; i < m; i += 256) gsum[i] = group_sum(b + i);
        __syncthreads();
        if (t == 0)
            for (int i = 0; i < m; ++i) total = total + gsum[i];
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
    // the group
    double cum = 0.0;
    int64_t gsel = -1;
    for (int64_t b = 0; b < ngroups; b += GT) {
        const int m = int(min<int64_t>(GT, ngroups - b));
        __syncthreads();
        for (int i = t; i < m; i += 256) gsum[i] = group_sum(b + i);
        __syncthreads();
        if (t == 0) {
            for (int i = 0; i < m && gsel < 0; ++i) {
                const int64_t g = b + i;
                if (g == ngroups - 1 || cum + gsum[i] > target)
                    gsel = g;
                else
                    cum = cum + gsum[i];
            }
        }
    }
    // the chunk inside the group
    if (t == 0) {
        const int64_t q0 = gsel * INIT_GROUP, q1 = min<int64_t>(nchunks, q0 + INIT_GROUP);
        int64_t qsel = q1 - 1;
        for (int64_t q = q0; q < q1; ++q) {
            const double c = chunk_sums[q];
            if (q == q1 - 1 || cum + c > target) {
                qsel = q;
                break;
            }
            cum = cum + c;
        }
        sh_sel = qsel;
        sh_cum = cum;
    }
    __syncthreads();
    const int64_t qsel = sh_sel;
    cum = sh_cum;
    // block sums of the chunk: thread j sums rows INIT_SUB * j, ... in order
    const int64_t r0 = qsel * OBJ_CHUNK;
    const int rows = int(min<int64_t>(OBJ_CHUNK, n - r0));
    const int nb = (rows + INIT_SUB - 1) / INIT_SUB;
    if (t < nb) {
        const int e = min(rows, (t + 1) * INIT_SUB);
        double sb = 0.0;
        for (int r = t * INIT_SUB; r < e; ++r) sb = sb + D2[r0 + r];
        bsum[t] = sb;
    }
    __syncthreads();
    if (t == 0) {
        int jsel = nb - 1;
        for (int j = 0; j < nb; ++j) {
            if (j == nb - 1 || cum + bsum[j] > target) {
                jsel = j;
                break;
            }
            cum = cum + bsum[j];
        }
        int64_t idx = -1, last_pos = -1;
        double sr = cum;
        const int rb = jsel * INIT_SUB, re = min(rows, rb + INIT_SUB);
        for (int r = rb; r < re; ++r) {
            const double w = D2[r0 + r];
            if (w > 0.0) last_pos = r0 + r;
            sr = sr + w;
            if (w > 0.0 && sr > target) {
                idx = r0 + r;
                break;
            }
        }
        if (idx < 0 && last_pos < 0) {
            for (int r = rows - 1; r >= 0; --r)
                if (D2[r0 + r] > 0.0) {
                    last_pos = r0 + r;
                    break;
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

// ---------------------------------------------------------------------------
// CSR rows.  The kernels below take rows whose column indices (minus `base`)
// increase strictly within each row; init_centroids_csr merges unsorted or repeated
// columns first.  A row then stands for its dense expansion: x_f = F(0) + v at
// a stored column f (the expansion adds each stored value to a zero), 0
// elsewhere.
// ---------------------------------------------------------------------------
constexpr int INIT_CSR_FT = 1024;  // centroid features staged per pass

// init_d2_kernel on the dense expansion, without forming it: each thread walks
// its row's features in order (a missing feature contributes (0 - c_f)^2), so
// every sub, mul and add is the dense kernel's and D2 agrees bit for bit.
template <typename F>
__global__ __launch_bounds__(256) void init_d2_csr_kernel(const F* __restrict__ vals,
                                                          const int64_t* __restrict__ cols,
                                                          const int64_t* __restrict__ offs, int64_t base,
                                                          int64_t n, int d, const F* __restrict__ c,
                                                          double* __restrict__ D2, int first) {
    __shared__ double cs[INIT_CSR_FT];
    const int t = threadIdx.x;
    for (int64_t r0 = int64_t(blockIdx.x) * 256; r0 < n; r0 += int64_t(gridDim.x) * 256) {
        const int64_t r = r0 + t;
        const bool ok = r < n;
        int64_t p = ok ? offs[r] - base : 0;
        const int64_t e = ok ? offs[r + 1] - base : 0;
        int64_t nc = p < e ? cols[p] - base : INT64_MAX;  // next stored column
        double acc = 0.0;
        for (int f0 = 0; f0 < d; f0 += INIT_CSR_FT) {
            const int ft = min(INIT_CSR_FT, d - f0);
            __syncthreads();
            for (int q = t; q < ft; q += 256) cs[q] = double(c[f0 + q]);
            __syncthreads();
            if (ok) {
                for (int q = 0; q < ft; ++q) {
                    double xv = 0.0;
                    if (nc == int64_t(f0 + q)) {
                        xv = double(F(0) + vals[p]);
                        ++p;
                        nc = p < e ? cols[p] - base : INT64_MAX;
                    }
                    const double diff = xv - cs[q];
                    const double sq = diff * diff;
                    acc = acc + sq;
                }
            }
        }
        if (ok) D2[r] = first ? acc : fmin(D2[r], acc);
    }
}

// C[row_out + i] = dense expansion of row src_i for i < count: src_i = ids[i],
// or *chosen (count 1), or first + i; one workgroup per output row
template <typename F>
__global__ void init_expand_csr_rows_kernel(const F* __restrict__ vals, const int64_t* __restrict__ cols,
                                            const int64_t* __restrict__ offs, int64_t base, int d,
                                            const int64_t* __restrict__ ids, const int64_t* __restrict__ chosen,
                                            int64_t first, int64_t count, F* __restrict__ C, int64_t row_out) {
    for (int64_t i = blockIdx.x; i < count; i += gridDim.x) {
        const int64_t src = ids ? ids[i] : chosen ? chosen[0] : first + i;
        F* dst = C + (row_out + i) * int64_t(d);
        for (int f = threadIdx.x; f < d; f += blockDim.x) dst[f] = F(0);
        __syncthreads();  // zeros before the stored values (columns are distinct)
        const int64_t o0 = offs[src] - base, o1 = offs[src + 1] - base;
        for (int64_t p = o0 + threadIdx.x; p < o1; p += blockDim.x) dst[cols[p] - base] = F(0) + vals[p];
    }
}

}  // namespace km
