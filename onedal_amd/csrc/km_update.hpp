// km_update.hpp -- centroid update, objective and empty-cluster kernels.
//
// Replaces, on the oneDAL side:
//   count_clusters / count_empty_clusters   kernels_integral_dpc.cpp:45-149
//   partial_reduce_centroids                kernels_fp_impl.hpp:327-376
//   merge_reduce_centroids                  kernels_fp_impl.hpp:277-325
//   compute_objective_function              kernels_fp_impl.hpp:378-400
//   compute_squares (centroids)             kernels_fp_impl.hpp:402-414
//   find_candidates / fill_empty_clusters   empty_cluster_handling_dpc.cpp:26-150, 402-455
// The reference accumulates centroid sums in F with per-part copies walked by
// 128 subgroups; here rows are grouped by label (stable, ascending row index),
// each cluster's member list is cut into REDUCE_CHUNK chunks summed by one wave
// in double, and the chunk sums are merged in chunk order.  The order is fixed,
// so the result is identical run to run and identical to the oracle's
// restatement of it.
#pragma once

#include "km_common.hpp"

namespace km {

template <typename T>
__global__ void fill_kernel(T* __restrict__ p, int64_t count, T value) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < count;
         i += int64_t(gridDim.x) * blockDim.x)
        p[i] = value;
}

// copy rows into a zero-padded matrix with row stride ldx (for float4 row reads)
template <typename F>
__global__ void pad_rows_kernel(const F* __restrict__ X, int64_t n, int d, int64_t ldx,
                                F* __restrict__ out) {
    const int64_t total = n * ldx;
    for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total;
         e += int64_t(gridDim.x) * blockDim.x) {
        const int64_t r = e / ldx;
        const int c = int(e - r * ldx);
        out[e] = c < d ? X[r * d + c] : F(0);
    }
}

// ---------------------------------------------------------------------------
// row norms: ||x_i|| rounded up to float (used only by the error bound)
// ---------------------------------------------------------------------------
template <typename F>
__global__ __launch_bounds__(256) void row_norm_kernel(const F* __restrict__ X, int64_t n, int d,
                                                       float* __restrict__ norm_up) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * int64_t(blockDim.x) + threadIdx.x) >> 6;
    const int64_t nwaves = (int64_t(gridDim.x) * blockDim.x) >> 6;
    for (int64_t row = wave; row < n; row += nwaves) {
        double s = 0.0;
        for (int f = lane; f < d; f += 64) {
            const double v = double(X[row * d + f]);
            s = fma(v, v, s);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) {
            const double r = sqrt(s) * (1.0 + 1e-12);
            float f = float(r);
            if (double(f) < r) f = __uint_as_float(__float_as_uint(f) + 1u);  // next float up (f >= 0)
            norm_up[row] = f;
        }
    }
}

// ---------------------------------------------------------------------------
// centroid preparation for one assignment pass: ||c_j||^2 in F (from a double
// sum), max ||c_j|| (rounded up) and a zero-padded copy for the matrix-core
// kernel.  One wave per centroid.
// ---------------------------------------------------------------------------
template <typename F>
__global__ __launch_bounds__(256) void centroid_prep_kernel(const F* __restrict__ C, int k, int d,
                                                            int dpad, F* __restrict__ cn,
                                                            F* __restrict__ cpad,
                                                            unsigned long long* __restrict__ cmax_bits) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    for (int j = wave; j < k; j += nwaves) {
        double s = 0.0;
        for (int f = lane; f < d; f += 64) {
            const double v = double(C[int64_t(j) * d + f]);
            s = fma(v, v, s);
        }
        if (cpad)
            for (int f = lane; f < dpad; f += 64) cpad[int64_t(j) * dpad + f] = f < d ? C[int64_t(j) * d + f] : F(0);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) {
            cn[j] = F(s);
            const double r = sqrt(s) * (1.0 + 1e-12);
            atomicMax(cmax_bits, (unsigned long long)__double_as_longlong(r));
        }
    }
}

// ---------------------------------------------------------------------------
// label histogram (count_clusters)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void count_kernel(const int32_t* __restrict__ labels, int64_t n,
                                                    int k, int32_t* __restrict__ counts,
                                                    int use_lds) {
    extern __shared__ int32_t hist[];
    if (use_lds) {
        for (int j = threadIdx.x; j < k; j += blockDim.x) hist[j] = 0;
        __syncthreads();
    }
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int j = labels[i];
        if (use_lds)
            atomicAdd(&hist[j], 1);
        else
            atomicAdd(&counts[j], 1);
    }
    if (use_lds) {
        __syncthreads();
        for (int j = threadIdx.x; j < k; j += blockDim.x)
            if (hist[j]) atomicAdd(&counts[j], hist[j]);
    }
}

// ---------------------------------------------------------------------------
// single-workgroup scan over clusters: member offsets, chunk offsets and the
// ascending list of empty clusters.  meta[0] = #empty, meta[1] = #chunks.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void scan_clusters_kernel(const int32_t* __restrict__ counts,
                                                             int k, int64_t* __restrict__ offsets,
                                                             int32_t* __restrict__ chunk_base,
                                                             int32_t* __restrict__ empty_list,
                                                             int32_t* __restrict__ meta) {
    __shared__ int64_t s_off[1024];
    __shared__ int32_t s_chk[1024];
    __shared__ int32_t s_emp[1024];
    __shared__ int64_t carry_off;
    __shared__ int32_t carry_chk, carry_emp;
    const int t = threadIdx.x;
    if (t == 0) {
        carry_off = 0;
        carry_chk = 0;
        carry_emp = 0;
    }
    __syncthreads();
    for (int base = 0; base < k; base += 1024) {
        const int j = base + t;
        const int32_t c = j < k ? counts[j] : 0;
        s_off[t] = c;
        s_chk[t] = j < k ? (c + REDUCE_CHUNK - 1) / REDUCE_CHUNK : 0;
        s_emp[t] = (j < k && c == 0) ? 1 : 0;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int64_t a = 0;
            int32_t b = 0, e = 0;
            if (t >= off) {
                a = s_off[t - off];
                b = s_chk[t - off];
                e = s_emp[t - off];
            }
            __syncthreads();
            s_off[t] += a;
            s_chk[t] += b;
            s_emp[t] += e;
            __syncthreads();
        }
        if (j < k) {
            const int64_t incl_off = s_off[t];
            offsets[j] = carry_off + incl_off - c;
            chunk_base[j] = carry_chk + s_chk[t] - (c + REDUCE_CHUNK - 1) / REDUCE_CHUNK;
            if (c == 0) empty_list[carry_emp + s_emp[t] - 1] = j;
        }
        __syncthreads();
        if (t == 1023) {
            carry_off += s_off[1023];
            carry_chk += s_chk[1023];
            carry_emp += s_emp[1023];
        }
        __syncthreads();
    }
    if (t == 0) {
        offsets[k] = carry_off;
        chunk_base[k] = carry_chk;
        meta[0] = carry_emp;
        meta[1] = carry_chk;
    }
}

// ---------------------------------------------------------------------------
// Chunk sums.  One wave per chunk of REDUCE_CHUNK members of one cluster.
// G = reduce_groups(d) interleaved partial sums (member m of the chunk goes to
// partial m % G); lanes g*d + f (d <= 32) or f, f+64, ... (d > 32) own columns.
// CPL = columns per lane (d > 32 only).
// ---------------------------------------------------------------------------
template <typename F, int CPL>
__global__ __launch_bounds__(256) void reduce_chunks_kernel(
    const F* __restrict__ X, int d, int k, const int32_t* __restrict__ perm,
    const int64_t* __restrict__ offsets, const int32_t* __restrict__ chunk_base,
    const int32_t* __restrict__ meta, double* __restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    const int total = meta[1];
    const int G = reduce_groups(d);
    const bool small = d <= 32;
    const int g = small ? lane / d : 0;
    const int f0 = small ? lane % d : lane;
    const bool active = small ? (g < G) : true;
    for (int q = wave; q < total; q += nwaves) {
        // cluster owning chunk q: last j with chunk_base[j] <= q
        int lo = 0, hi = k;  // chunk_base[k] = total > q
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (chunk_base[mid] <= q) lo = mid; else hi = mid;
        }
        const int j = lo;
        const int64_t m_begin = offsets[j] + int64_t(q - chunk_base[j]) * REDUCE_CHUNK;
        const int64_t m_end = min(m_begin + REDUCE_CHUNK, offsets[j + 1]);
        double acc[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) acc[c] = 0.0;
        if (small) {
            for (int64_t m = m_begin; m < m_end; m += G) {
                const int64_t mm = m + g;
                if (active && mm < m_end) {
                    const int64_t row = perm[mm];
                    acc[0] = acc[0] + double(X[row * d + f0]);
                }
            }
            // partial g -> group 0, added in g order
            double tot = acc[0];
            for (int gg = 1; gg < G; ++gg) {
                const double o = __shfl(acc[0], gg * d + f0);
                tot = tot + o;
            }
            if (g == 0 && f0 < d) part[int64_t(q) * d + f0] = tot;
        } else {
            int64_t m = m_begin;
            for (; m + 8 <= m_end; m += 8) {
                int64_t r[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) r[u] = perm[m + u];
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const int f = f0 + 64 * c;
                    if (f < d) {
                        double v[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) v[u] = double(X[r[u] * d + f]);
#pragma unroll
                        for (int u = 0; u < 8; ++u) acc[c] = acc[c] + v[u];
                    }
                }
            }
            for (; m < m_end; ++m) {
                const int64_t r0 = perm[m];
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const int f = f0 + 64 * c;
                    if (f < d) acc[c] = acc[c] + double(X[r0 * d + f]);
                }
            }
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const int f = f0 + 64 * c;
                if (f < d) part[int64_t(q) * d + f] = acc[c];
            }
        }
    }
}

// Same chunk sums for 32 < d <= 512 with d % 4 == 0 and 16-byte aligned rows:
// a sub-wave of lpc lanes (a power of two, lpc * 4 * V >= d) per chunk, lane l
// owning the 4-vectors l, l + lpc, ..., V of them, so a member's row is read with
// 16-byte loads (eight members' loads in flight).  Each feature is still summed
// over the chunk's members in order, so the partials equal reduce_chunks_kernel's.
template <typename F, int V>
__global__ __launch_bounds__(256) void reduce_chunks_vec_kernel(
    const F* __restrict__ X, int d, int k, const int32_t* __restrict__ perm,
    const int64_t* __restrict__ offsets, const int32_t* __restrict__ chunk_base,
    const int32_t* __restrict__ meta, double* __restrict__ part, int lpc) {
    typedef F F4 __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63;
    const int sub = lane / lpc, sl = lane - sub * lpc;
    const int subs = 64 / lpc;
    const int64_t gsub = int64_t((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * subs + sub;
    const int64_t nsub = int64_t((gridDim.x * blockDim.x) >> 6) * subs;
    const int total = meta[1];
    const int d4 = d / 4;
    for (int64_t q = gsub; q < total; q += nsub) {
        // cluster owning chunk q: last j with chunk_base[j] <= q
        int lo = 0, hi = k;  // chunk_base[k] = total > q
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (chunk_base[mid] <= q) lo = mid; else hi = mid;
        }
        const int j = lo;
        const int64_t m_begin = offsets[j] + int64_t(q - chunk_base[j]) * REDUCE_CHUNK;
        const int64_t m_end = min(m_begin + REDUCE_CHUNK, offsets[j + 1]);
        double acc[V][4];
#pragma unroll
        for (int v = 0; v < V; ++v)
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[v][e] = 0.0;
        int64_t m = m_begin;
        // member ids one batch ahead of the row loads (the row loads of a batch
        // then wait on one memory latency, not two)
        int32_t r[8];
        if (m + 8 <= m_end) {
#pragma unroll
            for (int u = 0; u < 8; ++u) r[u] = perm[m + u];
        }
        for (; m + 8 <= m_end; m += 8) {
            int32_t rn[8];
            const bool more = m + 16 <= m_end;
#pragma unroll
            for (int u = 0; u < 8; ++u) rn[u] = more ? perm[m + 8 + u] : 0;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const int c4 = sl + lpc * v;
                if (c4 < d4) {
                    F4 xv[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        xv[u] = *reinterpret_cast<const F4*>(X + int64_t(r[u]) * d + 4 * c4);
#pragma unroll
                    for (int u = 0; u < 8; ++u)
#pragma unroll
                        for (int e = 0; e < 4; ++e) acc[v][e] = acc[v][e] + double(xv[u][e]);
                }
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) r[u] = rn[u];
        }
        for (; m < m_end; ++m) {
            const int64_t r0 = perm[m];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const int c4 = sl + lpc * v;
                if (c4 < d4) {
                    const F4 x4 = *reinterpret_cast<const F4*>(X + r0 * d + 4 * c4);
#pragma unroll
                    for (int e = 0; e < 4; ++e) acc[v][e] = acc[v][e] + double(x4[e]);
                }
            }
        }
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int c4 = sl + lpc * v;
            if (c4 < d4) {
                double* o = part + int64_t(q) * d + 4 * c4;
#pragma unroll
                for (int e = 0; e < 4; ++e) o[e] = acc[v][e];
            }
        }
    }
}


// ---------------------------------------------------------------------------
// merge chunk sums per cluster (chunk order) and divide: C_j = (F)(S1_j / n_j).
// Empty clusters are left for fill_empty_kernel.
// ---------------------------------------------------------------------------
template <typename F>
__global__ __launch_bounds__(256) void finalize_centroids_kernel(
    const double* __restrict__ part, const int32_t* __restrict__ counts,
    const int32_t* __restrict__ chunk_base, int k, int d, F* __restrict__ Cnext) {
    const int64_t total = int64_t(k) * d;
    for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total;
         e += int64_t(gridDim.x) * blockDim.x) {
        const int j = int(e / d), f = int(e % d);
        const int32_t cnt = counts[j];
        if (cnt == 0) continue;
        double s = 0.0;
        int q = chunk_base[j];
        const int q1 = chunk_base[j + 1];
        // loads issued ahead, adds kept in chunk order
        for (; q + 8 <= q1; q += 8) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = part[int64_t(q + u) * d + f];
#pragma unroll
            for (int u = 0; u < 8; ++u) s = s + v[u];
        }
        for (; q < q1; ++q) s = s + part[int64_t(q) * d + f];
        Cnext[e] = F(s / double(cnt));
    }
}

// ---------------------------------------------------------------------------
// objective: d64 per row against its label's centroid.  Order (fixed): rows in
// chunks of OBJ_CHUNK; thread t of a chunk sums rows t, t+256, t+512, ... in
// order; the 256 partials are combined by the tree p[t] += p[t+s], s = 128..1;
// chunk sums are added in chunk order by sum_sequential_kernel.
// ---------------------------------------------------------------------------
// Rows reach the threads through LDS, one wave at a time: wave w of a 256-row
// group owns rows w*64 .. w*64+63 (thread t = row t of the group); per 128-byte
// slice of the features (OBJ_SLICE<F> elements) the wave stages its 64 rows with
// contiguous loads into its own LDS rows, and each lane adds its row's squared
// differences slice after slice, reading its centroid's slice directly (16-byte
// loads; the centroids stay in cache), so d64 keeps the feature order.  The next
// slice's row loads are issued before the current slice is summed.
template <typename F>
struct obj_slice {
    static constexpr int FT = 128 / int(sizeof(F));  // elements per slice
    static constexpr int ST = FT + 1;                 // LDS row stride (bank spread)
};
template <typename F>
__host__ inline size_t objective_lds_bytes() {
    return size_t(256) * obj_slice<F>::ST * sizeof(F);
}

template <typename F>
__global__ __launch_bounds__(256) void objective_chunks_kernel(
    const F* __restrict__ X, int64_t n, int d, const F* __restrict__ C,
    const int32_t* __restrict__ labels, double* __restrict__ d64_out,
    double* __restrict__ chunk_sums) {
    constexpr int FT = obj_slice<F>::FT, ST = obj_slice<F>::ST;
    constexpr int V = 16 / int(sizeof(F));            // elements per 16-byte load
    constexpr int PPR = FT / V;                       // 16-byte parts per row slice
    __shared__ double part[256];
    extern __shared__ __align__(16) unsigned char obj_smem[];
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    F* xs = reinterpret_cast<F*>(obj_smem) + wave * 64 * ST;  // this wave's [64][ST]
    const bool vec = (d % V) == 0 && (reinterpret_cast<uintptr_t>(X) % 16) == 0 &&
                     (reinterpret_cast<uintptr_t>(C) % 16) == 0;
    typedef F FV __attribute__((ext_vector_type(V)));
    const int64_t nchunks = (n + OBJ_CHUNK - 1) / OBJ_CHUNK;
    for (int64_t q = blockIdx.x; q < nchunks; q += gridDim.x) {
        const int64_t r0 = q * OBJ_CHUNK;
        const int rows = int(min<int64_t>(OBJ_CHUNK, n - r0));
        double p = 0.0;
        // thread t owns chunk rows t, t+256, ... in increasing order
        for (int g0 = 0; g0 < rows; g0 += 256) {
            const int nr = min(256, rows - g0);
            const int64_t wb = r0 + g0 + wave * 64;        // the wave's first row
            const int wn = max(0, min(64, nr - wave * 64));  // its rows in this group
            const bool mine = lane < wn;
            const F* c = C + int64_t(mine ? labels[wb + lane] : 0) * d;
            double acc = 0.0;
            if (vec) {
                // the wave's row slice f0: parts e = lane + 64 u, (row, part) = (e / PPR, e % PPR)
                FV xv[PPR];
                auto load = [&](int f0) {
#pragma unroll
                    for (int u = 0; u < PPR; ++u) {
                        const int e = lane + 64 * u, r = e / PPR, f = (e - r * PPR) * V;
                        xv[u] = (r < wn && f0 + f < d) ? *reinterpret_cast<const FV*>(X + (wb + r) * d + f0 + f)
                                                       : FV{};
                    }
                };
                load(0);
                for (int f0 = 0; f0 < d; f0 += FT) {
                    const int ft = min(FT, d - f0);
                    __builtin_amdgcn_wave_barrier();  // the previous slice's readers are done
#pragma unroll
                    for (int u = 0; u < PPR; ++u) {
                        const int e = lane + 64 * u, r = e / PPR, f = (e - r * PPR) * V;
#pragma unroll
                        for (int v = 0; v < V; ++v) xs[r * ST + f + v] = xv[u][v];
                    }
                    __builtin_amdgcn_wave_barrier();
                    if (f0 + FT < d) load(f0 + FT);
                    if (mine) {
                        for (int f = 0; f < ft; f += V) {
                            const FV cv = *reinterpret_cast<const FV*>(c + f0 + f);
#pragma unroll
                            for (int v = 0; v < V; ++v) {
                                const double diff = double(xs[lane * ST + f + v]) - double(cv[v]);
                                const double sq = diff * diff;
                                acc = acc + sq;
                            }
                        }
                    }
                }
            } else {
                for (int f0 = 0; f0 < d; f0 += FT) {
                    const int ft = min(FT, d - f0);
                    __builtin_amdgcn_wave_barrier();
                    for (int e = lane; e < wn * ft; e += 64) {
                        const int r = e / ft, f = e - r * ft;
                        xs[r * ST + f] = X[(wb + r) * d + f0 + f];
                    }
                    __builtin_amdgcn_wave_barrier();
                    if (mine) {
                        for (int f = 0; f < ft; ++f) {
                            const double diff = double(xs[lane * ST + f]) - double(c[f0 + f]);
                            const double sq = diff * diff;
                            acc = acc + sq;
                        }
                    }
                }
            }
            if (mine) {
                if (d64_out) d64_out[wb + lane] = acc;
                p = p + acc;
            }
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

__global__ void sum_sequential_kernel(const double* __restrict__ v, int64_t count,
                                      double* __restrict__ out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double s = 0.0;
        int64_t i = 0;
        for (; i + 16 <= count; i += 16) {
            double w[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) w[u] = v[i + u];
#pragma unroll
            for (int u = 0; u < 16; ++u) s = s + w[u];
        }
        for (; i < count; ++i) s = s + v[i];
        *out = s;
    }
}

// *acc += v[0] + ... + v[count-1], added one at a time in order: a sum over row
// blocks processed one after another equals sum_sequential_kernel over all
// chunks at once (acc starts at 0, and 0 + v0 = v0).
__global__ void sum_sequential_into_kernel(const double* __restrict__ v, int64_t count,
                                           double* __restrict__ acc) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double s = *acc;
        for (int64_t i = 0; i < count; ++i) s = s + v[i];
        *acc = s;
    }
}

// ---------------------------------------------------------------------------
// empty clusters: radix select of the m largest d64 (keys are the raw bits of
// non-negative doubles, which order like the values).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void digit_hist_kernel(const double* __restrict__ d64, int64_t n,
                                                         unsigned long long prefix,
                                                         unsigned long long mask, int shift,
                                                         unsigned int* __restrict__ hist) {
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const unsigned long long key = (unsigned long long)__double_as_longlong(d64[i]);
        if ((key & mask) == prefix) atomicAdd(&h[(key >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

__global__ __launch_bounds__(256) void collect_candidates_kernel(
    const double* __restrict__ d64, int64_t n, unsigned long long thr_key,
    int32_t* __restrict__ gt_list, int32_t* __restrict__ eq_list, int32_t* __restrict__ counters) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const unsigned long long key = (unsigned long long)__double_as_longlong(d64[i]);
        if (key > thr_key) gt_list[atomicAdd(&counters[0], 1)] = int32_t(i);
        else if (key == thr_key) eq_list[atomicAdd(&counters[1], 1)] = int32_t(i);
    }
}

__global__ void gather_f64_kernel(const double* __restrict__ src, const int32_t* __restrict__ idx,
                                  int64_t count, double* __restrict__ out) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < count;
         i += int64_t(gridDim.x) * blockDim.x)
        out[i] = src[idx[i]];
}

template <typename F>
__global__ __launch_bounds__(256) void fill_empty_kernel(const F* __restrict__ X, int d,
                                                         const int32_t* __restrict__ empty_list,
                                                         const int32_t* __restrict__ cand, int m,
                                                         F* __restrict__ Cnext) {
    const int64_t total = int64_t(m) * d;
    for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total;
         e += int64_t(gridDim.x) * blockDim.x) {
        const int q = int(e / d), f = int(e % d);
        Cnext[int64_t(empty_list[q]) * d + f] = X[int64_t(cand[q]) * d + f];
    }
}

}  // namespace km
