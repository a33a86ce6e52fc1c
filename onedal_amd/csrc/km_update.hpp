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
// Empty clusters are left for the ec_* kernels (refilled there).
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
        for (; q + 16 <= q1; q += 16) {
            double v[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) v[u] = part[int64_t(q + u) * d + f];
#pragma unroll
            for (int u = 0; u < 16; ++u) s = s + v[u];
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
// Thread t reads its rows directly, two rows (t + 256 j and t + 256 (j + 1)) at a
// time so their loads overlap: one 128-byte line of each row and of its
// centroid per step (16-byte loads), each row's squares added in feature order;
// the two rows' d64 then enter the thread's sum in row order.  No LDS.
template <typename F>
__host__ inline size_t objective_lds_bytes() {
    return 0;
}

template <typename F>
__global__ __launch_bounds__(256) void objective_chunks_kernel(
    const F* __restrict__ X, int64_t n, int d, const F* __restrict__ C,
    const int32_t* __restrict__ labels, double* __restrict__ d64_out,
    double* __restrict__ chunk_sums, const int32_t* __restrict__ gate) {
    if (gate && gate[0] == 0) return;  // gated off (no empty clusters this iteration)
    constexpr int RPT = 2;                           // rows in flight per thread
    constexpr int LV = 128 / (4 * int(sizeof(F)));   // 4-vectors per 128-byte line
    __shared__ double part[256];
    const int t = threadIdx.x;
    const bool vec = (d % 4) == 0 && (reinterpret_cast<uintptr_t>(X) % (4 * sizeof(F))) == 0 &&
                     (reinterpret_cast<uintptr_t>(C) % (4 * sizeof(F))) == 0;
    typedef F F4 __attribute__((ext_vector_type(4)));
    const int64_t nchunks = (n + OBJ_CHUNK - 1) / OBJ_CHUNK;
    for (int64_t q = blockIdx.x; q < nchunks; q += gridDim.x) {
        const int64_t r0 = q * OBJ_CHUNK;
        const int rows = int(min<int64_t>(OBJ_CHUNK, n - r0));
        double p = 0.0;
        // thread t owns chunk rows t, t+256, ... and adds them in that order
        for (int r = t; r < rows; r += 256 * RPT) {
            double acc[RPT];
            const F* xs[RPT];
            const F* cs[RPT];
            bool ok[RPT];
#pragma unroll
            for (int u = 0; u < RPT; ++u) {
                const int rr = r + 256 * u;
                ok[u] = rr < rows;
                const int64_t row = r0 + (ok[u] ? rr : r);
                xs[u] = X + row * d;
                cs[u] = C + int64_t(labels[row]) * d;
                acc[u] = 0.0;
            }
            int f = 0;
            if (vec) {
                for (; f + 4 * LV <= d; f += 4 * LV) {
                    F4 xv[RPT][LV], cv[RPT][LV];
#pragma unroll
                    for (int u = 0; u < RPT; ++u)
#pragma unroll
                        for (int v = 0; v < LV; ++v) {
                            xv[u][v] = *reinterpret_cast<const F4*>(xs[u] + f + 4 * v);
                            cv[u][v] = *reinterpret_cast<const F4*>(cs[u] + f + 4 * v);
                        }
#pragma unroll
                    for (int u = 0; u < RPT; ++u)
#pragma unroll
                        for (int v = 0; v < LV; ++v)
#pragma unroll
                            for (int e = 0; e < 4; ++e) {
                                const double diff = double(xv[u][v][e]) - double(cv[u][v][e]);
                                const double sq = diff * diff;
                                acc[u] = acc[u] + sq;
                            }
                }
            }
            for (; f < d; ++f)
#pragma unroll
                for (int u = 0; u < RPT; ++u) {
                    const double diff = double(xs[u][f]) - double(cs[u][f]);
                    const double sq = diff * diff;
                    acc[u] = acc[u] + sq;
                }
#pragma unroll
            for (int u = 0; u < RPT; ++u)
                if (ok[u]) {
                    if (d64_out) d64_out[r0 + r + 256 * u] = acc[u];
                    p = p + acc[u];
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
// Empty clusters on the device, as a fixed sequence of ordinary launches per
// iteration, each of which ends at once when meta[0] = 0 (no empty cluster):
//   ec_begin_kernel     adds the iteration's re-score counters into acc[0..2];
//                       with m = meta[0] > 0 empty clusters, zeroes the
//                       histograms;
//   ec_hist_kernel x 8  radix select of the m-th largest d64, 8 bits a pass from
//                       the top (keys: the raw bits of the non-negative doubles,
//                       which order like the values); every workgroup derives
//                       the digits of the earlier passes from their histograms;
//   ec_count_kernel,    rows above the threshold (gt) and the first `remaining`
//   ec_collect_kernel   rows equal to it (eq), collected in row order (workgroup
//                       w owns a contiguous row range; block scans give in-order
//                       ranks);
//   ec_order_kernel     gt rows ordered by (d64 descending, row ascending), eq
//                       rows after them: pick e is the row with the e-th largest
//                       d64, ties to the lowest row;
//   ec_finish_kernel    pick e fills the e-th empty cluster (ascending index) of
//                       Cnext (dense rows; sparse rows are filled by
//                       csr_fill_empty_kernel from cand), the objective (when
//                       given) drops by the picks' d64 in pick order, and acc[3]
//                       counts the iteration.
// The d64 of every row against its label's centroid comes in from the
// objective pass (gated on the same count).
// ---------------------------------------------------------------------------
constexpr int EC_PASSES = 8;

// threshold state after `passes` radix passes: key prefix, its mask, and how
// many rows equal to the final key are still to be taken
__device__ __forceinline__ void ec_state(const unsigned int* __restrict__ hist, int passes, int m,
                                         unsigned long long& prefix, unsigned long long& mask,
                                         long long& rem) {
    prefix = 0ull;
    mask = 0ull;
    rem = m;
    for (int p = 0; p < passes; ++p) {
        const int shift = 56 - 8 * p;
        long long cum = 0;
        int digit = 0;
        for (int dg = 255; dg >= 0; --dg) {
            const long long c = hist[p * 256 + dg];
            if (cum + c >= rem) {
                digit = dg;
                rem -= cum;
                break;
            }
            cum += c;
        }
        prefix |= (unsigned long long)digit << shift;
        mask |= 0xFFull << shift;
    }
}

// the workgroup's copy of the state (thread 0 computes it)
__device__ __forceinline__ void ec_state_shared(const unsigned int* __restrict__ hist, int passes, int m,
                                                unsigned long long& prefix, unsigned long long& mask,
                                                long long& rem) {
    __shared__ unsigned long long s_prefix, s_mask;
    __shared__ long long s_rem;
    if (threadIdx.x == 0) ec_state(hist, passes, m, s_prefix, s_mask, s_rem);
    __syncthreads();
    prefix = s_prefix;
    mask = s_mask;
    rem = s_rem;
}

__device__ __forceinline__ unsigned long long ec_key(double v) {
    return (unsigned long long)__double_as_longlong(v);
}

__global__ __launch_bounds__(256) void ec_begin_kernel(const int32_t* __restrict__ meta,
                                                       const int32_t* __restrict__ counters,
                                                       unsigned long long* __restrict__ acc,
                                                       unsigned int* __restrict__ hist) {
    if (threadIdx.x == 0 && counters) {
        acc[0] += (unsigned long long)counters[0];
        acc[1] += (unsigned long long)counters[1];
        acc[2] += (unsigned long long)counters[2];
    }
    if (meta[0] <= 0) return;
    for (int i = threadIdx.x; i < EC_PASSES * 256; i += blockDim.x) hist[i] = 0u;
}

__global__ __launch_bounds__(256) void ec_hist_kernel(const double* __restrict__ d64, int64_t n,
                                                      const int32_t* __restrict__ meta,
                                                      unsigned int* __restrict__ hist, int pass) {
    const int m = meta[0];
    if (m <= 0) return;
    __shared__ unsigned int h[256];
    unsigned long long prefix, mask;
    long long rem;
    ec_state_shared(hist, pass, m, prefix, mask, rem);
    const int shift = 56 - 8 * pass;
    h[threadIdx.x] = 0u;
    __syncthreads();
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const unsigned long long key = ec_key(d64[i]);
        if ((key & mask) == prefix) atomicAdd(&h[(key >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[pass * 256 + threadIdx.x], h[threadIdx.x]);
}

// workgroup w's contiguous row range
__device__ __forceinline__ void ec_range(int64_t n, int64_t& lo, int64_t& hi) {
    const int64_t span = (n + gridDim.x - 1) / gridDim.x;
    lo = min<int64_t>(n, int64_t(blockIdx.x) * span);
    hi = min<int64_t>(n, lo + span);
}

__global__ __launch_bounds__(256) void ec_count_kernel(const double* __restrict__ d64, int64_t n,
                                                       const int32_t* __restrict__ meta,
                                                       const unsigned int* __restrict__ hist,
                                                       int32_t* __restrict__ wcnt) {
    const int m = meta[0];
    if (m <= 0) return;
    __shared__ int32_t sa[256], sb[256];
    const int t = threadIdx.x;
    unsigned long long prefix, mask;
    long long rem;
    ec_state_shared(hist, EC_PASSES, m, prefix, mask, rem);
    int64_t lo, hi;
    ec_range(n, lo, hi);
    int cg_ = 0, ce = 0;
    for (int64_t i = lo + t; i < hi; i += 256) {
        const unsigned long long key = ec_key(d64[i]);
        cg_ += key > prefix;
        ce += key == prefix;
    }
    sa[t] = cg_;
    sb[t] = ce;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if (t < s2) {
            sa[t] += sa[t + s2];
            sb[t] += sb[t + s2];
        }
        __syncthreads();
    }
    if (t == 0) {
        wcnt[2 * blockIdx.x] = sa[0];
        wcnt[2 * blockIdx.x + 1] = sb[0];
    }
}

__global__ __launch_bounds__(256) void ec_collect_kernel(const double* __restrict__ d64, int64_t n,
                                                         const int32_t* __restrict__ meta,
                                                         const unsigned int* __restrict__ hist,
                                                         const int32_t* __restrict__ wcnt,
                                                         int32_t* __restrict__ gt_list,
                                                         int32_t* __restrict__ eq_list) {
    const int m = meta[0];
    if (m <= 0) return;
    __shared__ int32_t sa[256], sb[256];
    const int t = threadIdx.x;
    unsigned long long prefix, mask;
    long long rem;
    ec_state_shared(hist, EC_PASSES, m, prefix, mask, rem);
    int64_t lo, hi;
    ec_range(n, lo, hi);
    // this workgroup's first gt and eq positions: the counts of the earlier ones
    long long bg = 0, be = 0;
    for (int w = t; w < int(blockIdx.x); w += 256) {
        bg += wcnt[2 * w];
        be += wcnt[2 * w + 1];
    }
    sa[t] = int(bg);
    sb[t] = int(be);
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if (t < s2) {
            sa[t] += sa[t + s2];
            sb[t] += sb[t + s2];
        }
        __syncthreads();
    }
    bg = sa[0];
    be = sb[0];
    __syncthreads();
    for (int64_t tile = lo; tile < hi; tile += 256) {
        const int64_t i = tile + t;
        bool fg = false, fe = false;
        if (i < hi) {
            const unsigned long long key = ec_key(d64[i]);
            fg = key > prefix;
            fe = key == prefix;
        }
        sa[t] = fg;
        sb[t] = fe;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {  // inclusive scans
            const int a = t >= o ? sa[t - o] : 0, b = t >= o ? sb[t - o] : 0;
            __syncthreads();
            sa[t] += a;
            sb[t] += b;
            __syncthreads();
        }
        if (fg) gt_list[bg + sa[t] - 1] = int32_t(i);
        if (fe && be + sb[t] - 1 < rem) eq_list[be + sb[t] - 1] = int32_t(i);
        bg += sa[255];
        be += sb[255];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void ec_order_kernel(const double* __restrict__ d64,
                                                       const int32_t* __restrict__ meta,
                                                       const unsigned int* __restrict__ hist,
                                                       const int32_t* __restrict__ gt_list,
                                                       const int32_t* __restrict__ eq_list,
                                                       int32_t* __restrict__ cand, double* __restrict__ cand_d) {
    const int m = meta[0];
    if (m <= 0) return;
    unsigned long long prefix, mask;
    long long rem;
    ec_state_shared(hist, EC_PASSES, m, prefix, mask, rem);
    const long long ngt = m - rem;
    const int64_t gtid = int64_t(blockIdx.x) * 256 + threadIdx.x, gsize = int64_t(gridDim.x) * 256;
    for (long long i = gtid; i < ngt; i += gsize) {
        const int32_t ri = gt_list[i];
        const unsigned long long ki = ec_key(d64[ri]);
        long long rank = 0;
        for (long long j = 0; j < ngt; ++j) {
            const int32_t rj = gt_list[j];
            const unsigned long long kj = ec_key(d64[rj]);
            rank += (kj > ki) || (kj == ki && rj < ri);
        }
        cand[rank] = ri;
        cand_d[rank] = d64[ri];
    }
    for (long long e = gtid; e < rem; e += gsize) {
        cand[ngt + e] = eq_list[e];
        cand_d[ngt + e] = d64[eq_list[e]];
    }
}

template <typename F>
__global__ __launch_bounds__(256) void ec_finish_kernel(const F* __restrict__ X, int d,
                                                        const int32_t* __restrict__ meta,
                                                        const int32_t* __restrict__ empty_list,
                                                        const int32_t* __restrict__ cand,
                                                        const double* __restrict__ cand_d, F* __restrict__ Cnext,
                                                        double* __restrict__ obj,
                                                        unsigned long long* __restrict__ acc) {
    const int m = meta[0];
    if (m <= 0) return;
    const int64_t gtid = int64_t(blockIdx.x) * 256 + threadIdx.x, gsize = int64_t(gridDim.x) * 256;
    if (X)
        for (int64_t e = gtid; e < int64_t(m) * d; e += gsize) {
            const int q = int(e / d), f = int(e - int64_t(q) * d);
            Cnext[int64_t(empty_list[q]) * d + f] = X[int64_t(cand[q]) * d + f];
        }
    if (gtid == 0) {
        if (obj) {
            double o = *obj;
            for (int e = 0; e < m; ++e) o = o - cand_d[e];
            *obj = o;
        }
        acc[3] += 1ull;
    }
}

}  // namespace km
