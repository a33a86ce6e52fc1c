// km_assign.hpp -- fused distance + argmin kernels (assign_clusters).
//
// Replaces oneDAL's assign_clusters = oneMKL gemm (D = -2 X C^T, written to HBM)
// + select_min_distance + complete_closest_distances
// (cpp/oneapi/dal/algo/kmeans/backend/gpu/kernels_fp_impl.hpp:127-275, 416-439).
// Here the n x k distance matrix never leaves registers: each lane keeps a
// running top-2 (value, index) over the centroids it sees (a "slot"), and a row
// is settled by the certification rule below.  Rows the rule cannot settle go
// to the exact re-score kernels in km_refine.hpp.
//
// Certification.  v_j is the approximate value of ||c_j||^2 - 2 x.c_j and
// band(x) bounds 2*(|v_j - V_j| + |D_j - dist_j|) (km_common.hpp).  With
// T = min_j v_j + band(x), every centroid whose exact-form distance D_j can be
// minimal has v_j <= T.  A slot holds its two smallest v; entries it dropped
// are >= its second value.  So if every slot's second value exceeds T, the
// set S = {slot minima <= T} contains every possible minimiser: |S| == 1 settles
// the row, 2 <= |S| <= MAXC sends S to the candidate re-score, anything else
// (a slot with two entries inside the band) sends the row to the full re-score.
#pragma once

#include "km_common.hpp"

namespace km {

template <typename F>
struct top2 {
    F m1, m2;
    int i1, i2;
    __device__ __forceinline__ void init() {
        m1 = fp_traits<F>::inf;
        m2 = fp_traits<F>::inf;
        i1 = -1;
        i2 = -1;
    }
    __device__ __forceinline__ void push(F v, int j) {
        if (v < m1) {
            m2 = m1;
            i2 = i1;
            m1 = v;
            i1 = j;
        } else if (v < m2) {
            m2 = v;
            i2 = j;
        }
    }
};

// Settle one row from its slots (see the file comment).  X = ||x|| rounded up,
// C = max_j ||c_j|| rounded up.
template <typename F>
__device__ __forceinline__ void decide_row(int64_t row, const top2<F>* slots, int nslots,
                                           const band_consts& bc, double X, double C,
                                           int32_t* __restrict__ labels, const refine_lists& L) {
    F best = fp_traits<F>::inf;
    int bi = -1;
    for (int s = 0; s < nslots; ++s) {
        if (slots[s].m1 < best || (slots[s].m1 == best && slots[s].i1 >= 0 && slots[s].i1 < bi)) {
            best = slots[s].m1;
            bi = slots[s].i1;
        }
    }
    const double T = bi >= 0 ? band_threshold(bc, X, C, double(best)) : 0.0;
    bool cert = (bi >= 0) && (T < double(fp_traits<F>::inf));
    int cnt = 0;
    int cand[MAXC];
    for (int s = 0; s < nslots; ++s) {
        if (double(slots[s].m2) <= T) cert = false;
        if (slots[s].i1 >= 0 && double(slots[s].m1) <= T) {
            if (cnt < MAXC) cand[cnt] = slots[s].i1;
            cnt++;
        }
    }
    if (cert && cnt == 1) {
        labels[row] = bi;
    } else if (cert && cnt <= MAXC) {
        const int e = atomicAdd(&L.counters[0], 1);
        L.cand_rows[e] = int32_t(row);
        L.cand_count[e] = cnt;
        for (int c = 0; c < cnt; ++c) L.cand_ids[int64_t(e) * MAXC + c] = cand[c];
    } else {
        const int e = atomicAdd(&L.counters[1], 1);
        L.full_rows[e] = int32_t(row);
    }
}

// ---------------------------------------------------------------------------
// Generic tile kernel (any d, any k; F = float or double).  256 threads, 64
// rows x 64 centroids per step, K chunks of 16 features staged in LDS, 4 x 4
// register micro-tile per thread.  Each row gets 16 slots (one per centroid
// column group).  FP64 vector FMA runs at the same rate as FP64 MFMA on this
// chip, so this is also the FP64 production kernel.
// ---------------------------------------------------------------------------
template <typename F>
__global__ __launch_bounds__(256) void assign_generic_kernel(
    const F* __restrict__ X, int64_t n, int d, const F* __restrict__ C, int k,
    const F* __restrict__ cn, const float* __restrict__ xnorm_up, const double* __restrict__ cmax_p,
    band_consts bc, int32_t* __restrict__ labels, refine_lists L) {
    constexpr int TN = 64, TM = 64, KC = 16, PAD = 4;
    __shared__ __align__(16) F Xs[KC][TN + PAD];
    __shared__ __align__(16) F Cs[KC][TM + PAD];
    __shared__ top2<F> slots[TN][16];

    const int tid = threadIdx.x;
    const int tx = tid & 15;   // centroid group
    const int ty = tid >> 4;   // row group
    const double cmax = *cmax_p;
    const int64_t ntiles = (n + TN - 1) / TN;

    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t n0 = tile * TN;
        top2<F> st[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) st[i].init();

        for (int m0 = 0; m0 < k; m0 += TM) {
            F acc[4][4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = F(0);

            for (int k0 = 0; k0 < d; k0 += KC) {
                // stage X[n0:n0+64, k0:k0+16] and C[m0:m0+64, k0:k0+16] transposed
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int e = tid + r * 256;
                    const int f = e & 15, p = e >> 4;
                    const int64_t row = n0 + p;
                    const int col = k0 + f;
                    Xs[f][p] = (row < n && col < d) ? X[row * d + col] : F(0);
                    const int cj = m0 + p;
                    Cs[f][p] = (cj < k && col < d) ? C[int64_t(cj) * d + col] : F(0);
                }
                __syncthreads();
#pragma unroll
                for (int f = 0; f < KC; ++f) {
                    F a[4], b[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        a[i] = Xs[f][ty * 4 + i];
                        b[i] = Cs[f][tx * 4 + i];
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
                }
                __syncthreads();
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int cj = m0 + tx * 4 + j;
                if (cj < k) {
                    const F c2 = cn[cj];
#pragma unroll
                    for (int i = 0; i < 4; ++i) st[i].push(fma(F(-2), acc[i][j], c2), cj);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) slots[ty * 4 + i][tx] = st[i];
        __syncthreads();
        if (tid < TN) {
            const int64_t row = n0 + tid;
            if (row < n) {
                decide_row<F>(row, slots[tid], 16, bc, double(xnorm_up[row]), cmax, labels, L);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// FP32 matrix-core kernel: v_mfma_f32_32x32x2_f32, d <= DP (DP in {32,64,128,256}).
// Rows are read as float4 with row stride ldx (a multiple of 4 floats, 16-byte
// aligned base; columns d..ldx-1 must be zero).
//
// Workgroup = 4 waves; each wave owns NB blocks of 32 rows, held in VGPRs for
// the whole pass (B operand; the next tile's rows are prefetched into a second
// register set), and reads centroids from LDS in blocks of 32 (A operand).
// Lane l = 32*h + r: h selects the feature half [h*DP/2, (h+1)*DP/2), so MFMA
// step s multiplies feature s of half 0 and feature DP/2+s of half 1.  The
// 32x32 accumulator puts D[m][n] (m = centroid, n = row) at lane n + 32*h',
// register i, with m = 8*(i/4) + 4*h' + (i%4): a lane sees 16 of the 32
// centroids of a block for one row.  Each lane keeps two slots (even / odd
// centroid blocks), so a row has four slots.
//
// RES: all centroids stay in LDS for the whole launch (loaded once per
// workgroup).  Otherwise blocks of 32 centroids are double-buffered through
// LDS, the next block prefetched into registers while the current one runs.
// ---------------------------------------------------------------------------
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

template <int DP>
struct mfma_f32_layout {
    static constexpr int HALF = DP / 2;
    static constexpr int CST = DP + 4;  // LDS row stride in floats
};

// LDS bytes needed by assign_mfma_f32_kernel
__host__ inline size_t mfma_f32_lds_bytes(int dp, int k, bool res) {
    const int cst = dp + 4;
    if (res) {
        const int kp = (k + 31) / 32 * 32;
        return size_t(kp) * cst * 4 + size_t(kp) * 4;
    }
    return size_t(2) * 32 * cst * 4 + 2 * 32 * 4;
}

template <int DP, int NB, bool RES>
__global__ __launch_bounds__(256, DP >= 256 ? 1 : 2) void assign_mfma_f32_kernel(
    const float* __restrict__ X, int64_t ldx, int64_t n, int d, const float* __restrict__ Cpad, int k,
    const float* __restrict__ cn, const float* __restrict__ xnorm_up,
    const double* __restrict__ cmax_p, band_consts bc, int32_t* __restrict__ labels,
    refine_lists L) {
    constexpr int HALF = DP / 2;
    constexpr int CST = DP + 4;
    constexpr int ROWS_PER_WAVE = 32 * NB;
    constexpr int ROWS_PER_WG = 4 * ROWS_PER_WAVE;
    constexpr int F4_PER_ROW = DP / 4;
    extern __shared__ __align__(16) float smem[];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int r32 = lane & 31;
    const int h = lane >> 5;
    const double cmax = *cmax_p;
    const int mblocks = (k + 31) / 32;
    const int kp = mblocks * 32;
    const int64_t ntiles = (n + ROWS_PER_WG - 1) / ROWS_PER_WG;

    float* Cs = smem;                                   // RES: kp rows; else 2 x 32 rows
    float* Cn = smem + (RES ? kp : 64) * CST;           // RES: kp; else 2 x 32

    if (RES) {
        for (int e = tid; e < kp * F4_PER_ROW; e += 256) {
            const int row = e / F4_PER_ROW, c4 = e % F4_PER_ROW;
            floatx4 v = {0.f, 0.f, 0.f, 0.f};
            if (row < k) v = *reinterpret_cast<const floatx4*>(Cpad + int64_t(row) * DP + c4 * 4);
            *reinterpret_cast<floatx4*>(&Cs[row * CST + c4 * 4]) = v;
        }
        for (int j = tid; j < kp; j += 256) Cn[j] = j < k ? cn[j] : fp_traits<float>::inf;
        __syncthreads();
    }

    // streamed centroid block: each thread moves STG float4 of a 32 x DP block
    constexpr int STG = (32 * F4_PER_ROW + 255) / 256;
    floatx4 pre[STG];
    float pre_cn = 0.f;
    auto prefetch = [&](int mb) {
#pragma unroll
        for (int q = 0; q < STG; ++q) {
            const int e = tid + q * 256;
            const int row = e / F4_PER_ROW, c4 = e % F4_PER_ROW;
            const int cj = mb * 32 + row;
            floatx4 v = {0.f, 0.f, 0.f, 0.f};
            if (e < 32 * F4_PER_ROW && cj < k)
                v = *reinterpret_cast<const floatx4*>(Cpad + int64_t(cj) * DP + c4 * 4);
            pre[q] = v;
        }
        if (tid < 32) pre_cn = (mb * 32 + tid < k) ? cn[mb * 32 + tid] : fp_traits<float>::inf;
    };
    auto commit = [&](int b) {
#pragma unroll
        for (int q = 0; q < STG; ++q) {
            const int e = tid + q * 256;
            if (e < 32 * F4_PER_ROW) {
                const int row = e / F4_PER_ROW, c4 = e % F4_PER_ROW;
                *reinterpret_cast<floatx4*>(&Cs[(b * 32 + row) * CST + c4 * 4]) = pre[q];
            }
        }
        if (tid < 32) Cn[b * 32 + tid] = pre_cn;
    };

    // rows of a tile into registers
    auto load_rows = [&](int64_t tile, float (&dst)[NB][HALF]) {
        const int64_t wrow0 = tile * ROWS_PER_WG + int64_t(wave) * ROWS_PER_WAVE;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int64_t row = wrow0 + nb * 32 + r32;
            const bool ok = tile < ntiles && row < n;
            const float* src = X + (ok ? row : 0) * ldx;
#pragma unroll
            for (int s = 0; s < HALF; s += 4) {
                const int f = h * HALF + s;
                floatx4 v = {0.f, 0.f, 0.f, 0.f};
                if (ok && f < d) v = *reinterpret_cast<const floatx4*>(src + f);
                dst[nb][s] = v[0];
                dst[nb][s + 1] = v[1];
                dst[nb][s + 2] = v[2];
                dst[nb][s + 3] = v[3];
            }
        }
    };

    // prefetch the next tile's rows only when both register sets fit
    constexpr bool PREF = DP <= 64 && NB * HALF <= 64;
    float xr[NB][HALF];
    float xn[NB][HALF];
    int64_t tile = blockIdx.x;
    if (PREF) load_rows(tile, xn);

    for (; tile < ntiles; tile += gridDim.x) {
        if (PREF) {
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                for (int s = 0; s < HALF; ++s) xr[nb][s] = xn[nb][s];
            load_rows(tile + gridDim.x, xn);  // in flight during this tile
        } else {
            load_rows(tile, xr);
        }

        top2<float> st[NB][2];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            st[nb][0].init();
            st[nb][1].init();
        }

        // one block of 32 centroids from LDS rows [lrow0, lrow0+32), global ids mb*32.., slot sl
        auto block = [&](int lrow0, const float* cnb, int mb, int sl) {
            floatx16 acc[NB];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                for (int i = 0; i < 16; ++i) acc[nb][i] = 0.f;
            const float* arow = &Cs[(lrow0 + r32) * CST + h * HALF];
#pragma unroll
            for (int s = 0; s < HALF; s += 4) {
                const floatx4 a4 = *reinterpret_cast<const floatx4*>(arow + s);
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        acc[nb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a4[q], xr[nb][s + q],
                                                                       acc[nb], 0, 0, 0);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const floatx4 c4 = *reinterpret_cast<const floatx4*>(cnb + 8 * q + 4 * h);
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int cj = mb * 32 + 8 * q + 4 * h + t;
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        st[nb][sl].push(__builtin_fmaf(-2.f, acc[nb][4 * q + t], c4[t]), cj);
                }
            }
        };

        if (RES) {
            int mb = 0;
            for (; mb + 1 < mblocks; mb += 2) {
                block(mb * 32, Cn + mb * 32, mb, 0);
                block(mb * 32 + 32, Cn + mb * 32 + 32, mb + 1, 1);
            }
            if (mb < mblocks) block(mb * 32, Cn + mb * 32, mb, 0);
        } else {
            __syncthreads();
            prefetch(0);
            commit(0);
            __syncthreads();
            int mb = 0;
            for (; mb + 1 < mblocks; mb += 2) {
                prefetch(mb + 1);
                block(0, Cn, mb, 0);
                commit(1);
                __syncthreads();
                if (mb + 2 < mblocks) prefetch(mb + 2);
                block(32, Cn + 32, mb + 1, 1);
                if (mb + 2 < mblocks) commit(0);
                __syncthreads();
            }
            if (mb < mblocks) {
                block(0, Cn, mb, 0);
                __syncthreads();
            }
        }

        // settle rows: lane r32 (h = 0) merges its two slots with lane r32 + 32's two
        const int64_t wrow0 = tile * ROWS_PER_WG + int64_t(wave) * ROWS_PER_WAVE;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            top2<float> sl[4];
            sl[0] = st[nb][0];
            sl[1] = st[nb][1];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                sl[2 + q].m1 = __shfl_xor(st[nb][q].m1, 32);
                sl[2 + q].m2 = __shfl_xor(st[nb][q].m2, 32);
                sl[2 + q].i1 = __shfl_xor(st[nb][q].i1, 32);
                sl[2 + q].i2 = __shfl_xor(st[nb][q].i2, 32);
            }
            const int64_t row = wrow0 + nb * 32 + r32;
            if (h == 0 && row < n)
                decide_row<float>(row, sl, 4, bc, double(xnorm_up[row]), cmax, labels, L);
        }
    }
}

// ---------------------------------------------------------------------------
// Certified FP64 matrix-core assignment (double data, d <= DP <= 256).
// x.c_j runs on v_mfma_f64_16x16x4f64: a wave takes NB blocks of 16 rows, lane
// (r16, q) holds row r16's features [q DP/4, (q+1) DP/4) in registers and feeds
// them in steps s (feature q DP/4 + s on the 4-deep inner dimension), the same
// mapping on the centroid side from LDS.  Centroids stream through LDS in chunks
// of CB = 64 (zero-padded to DP features and to whole blocks of 16).  Lane (r16,
// q) ends a block with v_j = ||c_j||^2 - 2 x.c_j for centroids q, q+4, q+8 and
// q+12 of the block (the instruction's output rows in that lane) and keeps two
// top-2 slots (even and odd blocks); a row's eight slots
// (its four lanes) are settled by decide_row with the FP64 error band, as the
// FP32 matrix-core kernel does, so labels equal the exact form's.
// ---------------------------------------------------------------------------
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef double doublex2 __attribute__((ext_vector_type(2)));
constexpr int MFMA_F64_CB = 64;  // centroids per LDS chunk

__host__ __device__ constexpr int mfma_f64_cst(int dp) { return dp + 2; }  // LDS row stride (doubles)
__host__ inline size_t mfma_f64_lds_bytes(int dp) {
    return size_t(MFMA_F64_CB) * mfma_f64_cst(dp) * 8 + MFMA_F64_CB * 8;
}

template <int DP, int NB>
__global__ __launch_bounds__(256, 2) void assign_mfma_f64_kernel(
    const double* __restrict__ X, int64_t ldx, int64_t n, int d, const double* __restrict__ Cpad, int k,
    const double* __restrict__ cn, const float* __restrict__ xnorm_up,
    const double* __restrict__ cmax_p, band_consts bc, int32_t* __restrict__ labels,
    refine_lists L) {
    constexpr int Q = DP / 4;           // features per lane (and MFMA steps per block)
    constexpr int CST = mfma_f64_cst(DP);
    constexpr int ROWS_PER_WAVE = 16 * NB;
    constexpr int ROWS_PER_WG = 4 * ROWS_PER_WAVE;
    extern __shared__ __align__(16) double dsm[];
    double* Cs = dsm;                          // [CB][CST]
    double* Cn = dsm + MFMA_F64_CB * CST;      // [CB]
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int r16 = lane & 15;
    const int q = lane >> 4;
    const double cmax = *cmax_p;
    const int64_t ntiles = (n + ROWS_PER_WG - 1) / ROWS_PER_WG;
    const int nchunks = (k + MFMA_F64_CB - 1) / MFMA_F64_CB;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t wrow0 = tile * ROWS_PER_WG + int64_t(wave) * ROWS_PER_WAVE;
        double xr[NB][Q];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int64_t row = wrow0 + nb * 16 + r16;
            const bool ok = row < n;
            const double* src = X + (ok ? row : 0) * ldx + q * Q;
#pragma unroll
            for (int s = 0; s < Q; s += 2) {
                doublex2 v = {0.0, 0.0};
                if (ok && q * Q + s < d) v = *reinterpret_cast<const doublex2*>(src + s);
                xr[nb][s] = v[0];
                xr[nb][s + 1] = v[1];
            }
        }
        top2<double> st[NB][2];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            st[nb][0].init();
            st[nb][1].init();
        }
        for (int ch = 0; ch < nchunks; ++ch) {
            const int c0 = ch * MFMA_F64_CB;
            __syncthreads();  // the previous chunk's readers are done
            for (int e = tid; e < MFMA_F64_CB * (DP / 2); e += 256) {
                const int r = e / (DP / 2), f = (e - r * (DP / 2)) * 2;
                doublex2 v = {0.0, 0.0};
                if (c0 + r < k) v = *reinterpret_cast<const doublex2*>(Cpad + int64_t(c0 + r) * DP + f);
                *reinterpret_cast<doublex2*>(Cs + r * CST + f) = v;
            }
            for (int r = tid; r < MFMA_F64_CB; r += 256) Cn[r] = c0 + r < k ? cn[c0 + r] : fp_traits<double>::inf;
            __syncthreads();
            const int blocks = min(MFMA_F64_CB, k - c0 + 15) / 16;
            for (int b = 0; b < blocks; ++b) {
                doublex4 acc[NB];
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) acc[nb] = doublex4{0.0, 0.0, 0.0, 0.0};
                const double* arow = Cs + (b * 16 + r16) * CST + q * Q;
#pragma unroll
                for (int s = 0; s < Q; s += 2) {
                    const doublex2 a2 = *reinterpret_cast<const doublex2*>(arow + s);
#pragma unroll
                    for (int u = 0; u < 2; ++u)
#pragma unroll
                        for (int nb = 0; nb < NB; ++nb)
                            acc[nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2[u], xr[nb][s + u], acc[nb], 0, 0, 0);
                }
                const int gb = (c0 >> 4) + b;  // global block index
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int lj = b * 16 + q + 4 * t;  // output row q + 4 t of the 16 x 16 block
                    const double cnj = Cn[lj];
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb) {
                        if (gb & 1)
                            st[nb][1].push(fma(-2.0, acc[nb][t], cnj), c0 + lj);
                        else
                            st[nb][0].push(fma(-2.0, acc[nb][t], cnj), c0 + lj);
                    }
                }
            }
        }
        // settle rows: lane (r16, 0) gathers the eight slots of its row
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            top2<double> sl[8];
#pragma unroll
            for (int w = 0; w < 4; ++w)
#pragma unroll
                for (int z = 0; z < 2; ++z) {
                    sl[2 * w + z].m1 = __shfl_xor(st[nb][z].m1, 16 * w);
                    sl[2 * w + z].m2 = __shfl_xor(st[nb][z].m2, 16 * w);
                    sl[2 * w + z].i1 = __shfl_xor(st[nb][z].i1, 16 * w);
                    sl[2 * w + z].i2 = __shfl_xor(st[nb][z].i2, 16 * w);
                }
            const int64_t row = wrow0 + nb * 16 + r16;
            if (q == 0 && row < n)
                decide_row<double>(row, sl, 8, bc, double(xnorm_up[row]), cmax, labels, L);
        }
    }
}

}  // namespace km
