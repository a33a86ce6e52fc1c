// km_batch.hpp -- batched Lloyd: thousands of independent instances per launch.
//
// The reference runs one instance per train() call
// (cpp/oneapi/dal/algo/kmeans/backend/gpu/train_kernel_lloyd_dense_dpc.cpp:76-207).
// Here every stage of one Lloyd iteration is one launch over all instances of a
// batch, so the iteration has no host synchronisation and no per-instance launch
// cost; the stages are the single-instance ones, and every instance's result is
// the single-instance result bit for bit (same arithmetic order):
//   exact-form assignment (features in order, ties to the lowest index)
//   -> label histogram -> per-instance scan (member offsets, chunk bases, empty
//   list) -> one stable radix sort of row ids by (instance, label) -> chunk sums in FP64
//   (REDUCE_CHUNK members, reduce_groups interleave) -> merge/divide
//   -> empty clusters refilled on the device with the farthest rows (largest d64,
//   ties to the lowest row) in ascending cluster order.
// Work items (row batches, chunks, centroid elements) of all instances form one
// index space; a workgroup or wave finds its instance by binary search in a
// prefix table.
#pragma once

#include "km_assign.hpp"
#include "km_common.hpp"

namespace km {

// One instance as the kernels see it.  Per-row arrays (labels, sorted row ids,
// d64) are concatenated over the batch (row0); centroid arrays too (c0, a
// multiple of 4 floats); cluster arrays (counts, empty lists) at k0, with the
// k + 1 entries of offsets and chunk bases at k0 + index; chunk partial sums at p0
// (room for n / REDUCE_CHUNK + k + 1 chunks of d doubles); objective chunk sums
// at o0.
struct batch_inst {
    const float* X;
    const float* C0;
    float* C_out;
    int32_t* labels_out;
    int64_t n;
    int64_t row0;
    int64_t c0;
    int64_t k0;
    int64_t p0;
    int64_t o0;
    int64_t h0;   // first of the instance's per-block label histograms (counting sort)
    int32_t d;
    int32_t k;
    int32_t index;
    int32_t vec;  // d % 4 == 0 and X 16-byte aligned
};

constexpr int BATCH_CNT_ROWS = 4096;      // rows per histogram workgroup
constexpr int BATCH_D64_ROWS = 256;       // rows per d64 workgroup
constexpr int BATCH_HIST_MAX = 4096;      // LDS histogram bins

// instance owning work item w: the last i with pre[i] <= w (pre[B] > w)
__device__ __forceinline__ int batch_find(const int64_t* __restrict__ pre, int B, int64_t w) {
    int lo = 0, hi = B;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= w) lo = mid; else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Exact-form assignment, tiled: a wave holds 64 / CT row groups of CT lanes; lane
// (group g, slot c) carries centroid j0 + c of a CT-centroid tile against RW rows,
// the centroid's features in registers (32 at a time, the next 32 loaded ahead),
// the rows staged in LDS (feature chunks of FT) and read as broadcasts within the
// group.  A workgroup (four waves) takes RB = 4 * RW * 64 / CT rows of one
// instance.  Each (row, centroid) sum runs over the features in order with sub,
// mul and add rounded separately, so every tile shape gives the same distances and
// labels; the shape only changes speed (the autotuner picks one per (d, k)
// bucket).  Scan semantics of select_min_distance (kernels_fp_impl.hpp:198-202):
// strict < in centroid order, a NaN distance to centroid 0 keeps label 0.
// The launch covers the instances sel[0..nsel) with pre the prefix of their
// workgroup counts.
// labels: concatenated labels (row0) or, with to_user, each instance's output;
// skeys (when labels are concatenated and skeys is not null): the row's sort key (index << lbits) |
// label, so one stable radix sort over the batch groups every instance's rows
// by label in row order.
// ---------------------------------------------------------------------------
// Tile variants 0..6 are exact-form VALU tiles; variant 7 is the certified FP32
// matrix-core kernel below (instances with d <= 128 and k <= 128).
constexpr int BATCH_TILES = 8;
constexpr int BATCH_TILE_MFMA = 7;
// (CT, RW) of each VALU tile variant
__host__ __device__ constexpr int batch_tile_ct(int v) {
    return v == 0 ? 64 : v == 1 ? 64 : v == 2 ? 32 : v == 3 ? 16 : v == 4 ? 8 : v == 5 ? 32 : 16;
}
__host__ __device__ constexpr int batch_tile_rw(int v) {
    return v == 0 ? 8 : v == 1 ? 4 : v == 2 ? 8 : v == 3 ? 8 : v == 4 ? 8 : v == 5 ? 4 : 4;
}
// rows per workgroup
__host__ __device__ constexpr int batch_tile_rows(int v) { return 4 * batch_tile_rw(v) * (64 / batch_tile_ct(v)); }

template <int CT, int RW>
__global__ __launch_bounds__(256) void batch_assign_kernel(const batch_inst* __restrict__ I,
                                                           const int32_t* __restrict__ sel, int nsel,
                                                           const int64_t* __restrict__ pre,
                                                           const float* __restrict__ C,
                                                           int32_t* __restrict__ labels,
                                                           uint32_t* __restrict__ skeys, int lbits,
                                                           int to_user) {
    constexpr int G = 64 / CT;    // row groups per wave
    constexpr int RWW = RW * G;   // rows per wave
    constexpr int RB = 4 * RWW;   // rows per workgroup
    constexpr int FT = 8192 / RB < 256 ? 8192 / RB : 256;  // staged features per row chunk
    constexpr int CQ = 8;         // 4-vectors per register chunk
    static_assert(FT % 4 == 0 && FT >= 4, "feature chunk");
    typedef float F4 __attribute__((ext_vector_type(4)));
    __shared__ __align__(16) float Xs[RB][FT];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int cl = lane % CT;
    const int rg = lane / CT;
    const int rbase = wave * RWW + rg * RW;  // first staged row of this lane
    const int64_t total = pre[nsel];
    for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
        const int si = batch_find(pre, nsel, b);
        const batch_inst in = I[sel[si]];
        const int d = in.d, k = in.k;
        const int64_t r0 = (b - pre[si]) * RB;
        const int nr = int(min<int64_t>(RB, in.n - r0));
        const float* Ci = C + in.c0;
        const bool vec = in.vec != 0;
        float best[RW];
        int bi[RW];
        bool nan0[RW];
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            best[r] = __builtin_inff();
            bi[r] = 0x7fffffff;
            nan0[r] = false;
        }
        // rows [r0, r0 + RB), features [f0, f0 + ft) into the LDS tile
        auto stage = [&](int f0, int ft) {
            __syncthreads();
            if (vec) {
                const int nv = ft / 4;
                for (int e = tid; e < RB * nv; e += 256) {
                    const int r = e / nv, q = e - r * nv;
                    F4 v = F4{0.f, 0.f, 0.f, 0.f};
                    if (r < nr) v = *reinterpret_cast<const F4*>(in.X + (r0 + r) * d + f0 + 4 * q);
                    *reinterpret_cast<F4*>(&Xs[r][4 * q]) = v;
                }
            } else {
                for (int e = tid; e < RB * ft; e += 256) {
                    const int r = e / ft, q = e - r * ft;
                    Xs[r][q] = r < nr ? in.X[(r0 + r) * d + f0 + q] : 0.f;
                }
            }
            __syncthreads();
        };
        // rows that fit one chunk are staged once for every centroid tile
        const bool one_chunk = d <= FT;
        if (one_chunk) stage(0, d);
        for (int j0 = 0; j0 < k; j0 += CT) {
            const int j = j0 + cl;
            const float* c = Ci + int64_t(j < k ? j : 0) * d;
            float acc[RW];
#pragma unroll
            for (int r = 0; r < RW; ++r) acc[r] = 0.f;
            for (int f0 = 0; f0 < d; f0 += FT) {
                const int ft = min(FT, d - f0);
                if (!one_chunk) stage(f0, ft);
                if (vec) {
                    const int nv = ft / 4;
                    const F4* c4 = reinterpret_cast<const F4*>(c + f0);
                    F4 cur[CQ], nxt[CQ];
#pragma unroll
                    for (int q = 0; q < CQ; ++q) cur[q] = q < nv ? c4[q] : F4{};
                    for (int u0 = 0; u0 < nv; u0 += CQ) {
#pragma unroll
                        for (int q = 0; q < CQ; ++q) nxt[q] = u0 + CQ + q < nv ? c4[u0 + CQ + q] : F4{};
#pragma unroll
                        for (int q = 0; q < CQ; ++q) {
                            if (u0 + q < nv) {
#pragma unroll
                                for (int r = 0; r < RW; ++r) {
                                    const F4 x4 = *reinterpret_cast<const F4*>(&Xs[rbase + r][4 * (u0 + q)]);
#pragma unroll
                                    for (int t = 0; t < 4; ++t) {
                                        const float diff = x4[t] - cur[q][t];
                                        const float sq = diff * diff;
                                        acc[r] = acc[r] + sq;
                                    }
                                }
                            }
                        }
#pragma unroll
                        for (int q = 0; q < CQ; ++q) cur[q] = nxt[q];
                    }
                } else {
                    for (int f = 0; f < ft; ++f) {
                        const float cf = c[f0 + f];
#pragma unroll
                        for (int r = 0; r < RW; ++r) {
                            const float diff = Xs[rbase + r][f] - cf;
                            const float sq = diff * diff;
                            acc[r] = acc[r] + sq;
                        }
                    }
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
            float bv = best[r];
            int bj = bi[r];
#pragma unroll
            for (int off = CT / 2; off > 0; off >>= 1) {
                const float ob = __shfl_xor(bv, off);
                const int oi = __shfl_xor(bj, off);
                if (ob < bv || (ob == bv && oi < bj)) {
                    bv = ob;
                    bj = oi;
                }
            }
            const bool n0 = __shfl(nan0[r] ? 1 : 0, rg * CT) != 0;
            const int rr = rbase + r;
            if (cl == 0 && rr < nr) {
                const int lab = (n0 || bj == 0x7fffffff) ? 0 : bj;
                if (to_user) {
                    in.labels_out[r0 + rr] = lab;
                } else {
                    labels[in.row0 + r0 + rr] = lab;
                    if (skeys) skeys[in.row0 + r0 + rr] = (uint32_t(in.index) << lbits) | uint32_t(lab);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Certified FP32 matrix-core assignment (tile variant 7) for instances with
// d <= DP (DP in {16, 32, 64, 128}) and k <= 128.  A work item is BATCH_MFMA_ROWS
// rows of one instance.  The item loads the instance's centroids into LDS once
// (zero-padded to DP features and to whole blocks of 32), with ||c_j||^2 (double
// sum rounded to float) and max ||c_j|| rounded up.  Then each 128-row tile puts
// a row in the registers of a lane pair (feature halves), runs
// v_mfma_f32_32x32x2f32 against every block of 32 centroids and keeps four top-2
// slots per row, as the single-instance FP32 kernel does (km_assign.hpp).  The
// settle uses that kernel's error band: v_j = ||c_j||^2 - 2 x.c_j is within the
// band of the exact-form distance minus ||x||^2, so every centroid that can be
// the exact-form argmin has v_j <= T.  Rows with one such centroid are settled.
// Rows with a few are re-scored exactly over those candidates (ties to the lowest
// index).  Every other row, and every row of an item whose centroids are not all
// finite, gets the exact scan of select_min_distance (kernels_fp_impl.hpp:198-202).
// So labels equal the exact-form kernel's.
// ---------------------------------------------------------------------------
constexpr int BATCH_MFMA_ROWS = 1024;  // rows per work item
constexpr int BATCH_MFMA_MAXK = 128;
__host__ __device__ constexpr size_t batch_mfma_lds_bytes(int dp) {
    return size_t(BATCH_MFMA_MAXK) * (dp + 4) * 4 + BATCH_MFMA_MAXK * 4 + 4 * 8;
}

// error-band constants of the FP32 matrix-core distance for d features (the
// host's make_band<float>)
__device__ __forceinline__ band_consts batch_band_f32(int d) {
    band_consts b;
    const double u = fp_traits<float>::u, ua = fp_traits<float>::u_acc;
    b.a_xc = 2.0 * gamma_m(double(d), ua) + 2.01 * u;
    b.a_cc = 2.01 * u + gamma_m(double(d), 2.220446049250313e-16);
    const double g = gamma_m(double(d + 2), u);
    b.g_canon = g / (1.0 - g);
    return b;
}

template <int DP>
__global__ __launch_bounds__(256, 2) void batch_assign_mfma_kernel(const batch_inst* __restrict__ I,
                                                                   const int32_t* __restrict__ sel, int nsel,
                                                                   const int64_t* __restrict__ pre,
                                                                   const float* __restrict__ C,
                                                                   int32_t* __restrict__ labels,
                                                                   uint32_t* __restrict__ skeys, int lbits,
                                                                   int to_user) {
    constexpr int HALF = DP / 2;
    constexpr int CST = DP + 4;  // LDS row stride in floats
    extern __shared__ __align__(16) float smem[];
    float* Cs = smem;                                                   // BATCH_MFMA_MAXK x CST
    float* Cn = smem + BATCH_MFMA_MAXK * CST;                           // BATCH_MFMA_MAXK
    double* wmax = reinterpret_cast<double*>(Cn + BATCH_MFMA_MAXK);     // one per wave
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int r32 = lane & 31;
    const int h = lane >> 5;
    const float finf = __builtin_inff();
    const int64_t total = pre[nsel];
    for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
        const int si = batch_find(pre, nsel, b);
        const batch_inst in = I[sel[si]];
        const int d = in.d, k = in.k;
        if (d > DP || k > BATCH_MFMA_MAXK) continue;  // never planned so; keeps LDS in bounds
        const int64_t rb = (b - pre[si]) * BATCH_MFMA_ROWS;
        const int64_t re = min<int64_t>(in.n, rb + BATCH_MFMA_ROWS);
        const float* Ci = C + in.c0;
        const int mblocks = (k + 31) / 32;
        const int kp = mblocks * 32;
        __syncthreads();  // the previous item's LDS reads are done
        for (int e = tid; e < kp * DP; e += 256) {
            const int j = e / DP, f = e - j * DP;
            Cs[j * CST + f] = (j < k && f < d) ? Ci[int64_t(j) * d + f] : 0.f;
        }
        __syncthreads();
        // ||c_j||^2 and the largest of them: one wave per centroid; a centroid
        // that is not finite sends every row of the item to the exact scan
        double wm = 0.0;
        bool wbad = false;
        for (int j = wave; j < kp; j += 4) {
            double sq = 0.0;
            for (int f = lane; f < DP; f += 64) {
                const double v = double(Cs[j * CST + f]);
                sq = fma(v, v, sq);
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off);
            if (lane == 0) Cn[j] = j < k ? float(sq) : finf;
            if (j < k) {
                if (!(sq < 1.0e300)) wbad = true;  // NaN or overflow
                wm = fmax(wm, sq);
            }
        }
        if (lane == 0) wmax[wave] = wbad ? -1.0 : wm;
        __syncthreads();
        const bool cbad = wmax[0] < 0.0 || wmax[1] < 0.0 || wmax[2] < 0.0 || wmax[3] < 0.0;
        const double cmax =
            sqrt(fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]))) * (1.0 + 1e-12);
        const band_consts bc = batch_band_f32(d);
        for (int64_t t0 = rb; t0 < re; t0 += 128) {
            const int64_t row = t0 + wave * 32 + r32;
            const bool ok = row < re;
            const float* xrow = in.X + (ok ? row : rb) * int64_t(d);
            float xr[HALF];
            if (in.vec) {
#pragma unroll
                for (int s4 = 0; s4 < HALF; s4 += 4) {
                    const int f = h * HALF + s4;
                    floatx4 v = {0.f, 0.f, 0.f, 0.f};
                    if (ok && f < d) v = *reinterpret_cast<const floatx4*>(xrow + f);
                    xr[s4] = v[0];
                    xr[s4 + 1] = v[1];
                    xr[s4 + 2] = v[2];
                    xr[s4 + 3] = v[3];
                }
            } else {
#pragma unroll
                for (int s1 = 0; s1 < HALF; ++s1) {
                    const int f = h * HALF + s1;
                    xr[s1] = (ok && f < d) ? xrow[f] : 0.f;
                }
            }
            // ||x|| rounded up (double sum over both halves)
            double xs = 0.0;
#pragma unroll
            for (int s1 = 0; s1 < HALF; ++s1) {
                const double v = double(xr[s1]);
                xs = fma(v, v, xs);
            }
            xs += __shfl_xor(xs, 32);
            const double Xup = sqrt(xs) * (1.0 + 1e-12);

            top2<float> st[2];
            st[0].init();
            st[1].init();
            // block mb into slot sl (even blocks slot 0, odd blocks slot 1; the
            // slot is a compile-time index, so the slots stay in registers)
            auto block = [&](int mb, top2<float>& slot) {
                floatx16 acc;
#pragma unroll
                for (int i = 0; i < 16; ++i) acc[i] = 0.f;
                const float* arow = &Cs[(mb * 32 + r32) * CST + h * HALF];
#pragma unroll
                for (int s4 = 0; s4 < HALF; s4 += 4) {
                    const floatx4 a4 = *reinterpret_cast<const floatx4*>(arow + s4);
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a4[q], xr[s4 + q], acc, 0, 0, 0);
                }
                const float* cnb = Cn + mb * 32;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const floatx4 c4 = *reinterpret_cast<const floatx4*>(cnb + 8 * q + 4 * h);
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                        slot.push(__builtin_fmaf(-2.f, acc[4 * q + t], c4[t]), mb * 32 + 8 * q + 4 * h + t);
                }
            };
            for (int mb = 0; mb < mblocks; mb += 2) {
                block(mb, st[0]);
                if (mb + 1 < mblocks) block(mb + 1, st[1]);
            }
            // lane r32 (h = 0) holds the row's four slots: its two and lane r32 + 32's
            top2<float> sl[4];
            sl[0] = st[0];
            sl[1] = st[1];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                sl[2 + q].m1 = __shfl_xor(st[q].m1, 32);
                sl[2 + q].m2 = __shfl_xor(st[q].m2, 32);
                sl[2 + q].i1 = __shfl_xor(st[q].i1, 32);
                sl[2 + q].i2 = __shfl_xor(st[q].i2, 32);
            }
            int lab = 0;
            bool need_scan = false;
            if (h == 0 && ok) {
                float best = finf;
                int bi = -1;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (sl[q].m1 < best || (sl[q].m1 == best && sl[q].i1 >= 0 && sl[q].i1 < bi)) {
                        best = sl[q].m1;
                        bi = sl[q].i1;
                    }
                }
                const double T = bi >= 0 ? band_threshold(bc, Xup, cmax, double(best)) : 0.0;
                bool cert = !cbad && bi >= 0 && bi < k && T < double(finf);
                int cnt = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (double(sl[q].m2) <= T) cert = false;
                    if (sl[q].i1 >= 0 && double(sl[q].m1) <= T) {
                        if (sl[q].i1 >= k) cert = false;
                        ++cnt;
                    }
                }
                if (cert && cnt == 1) {
                    lab = bi;
                } else if (cert) {
                    // exact form over the candidates (the slot minima inside T, in
                    // slot order; slots are visited by a fixed unrolled loop, so no
                    // candidate array is indexed at run time), ties to the lowest index
                    float bd = finf;
                    lab = k;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        if (sl[q].i1 >= 0 && double(sl[q].m1) <= T) {
                            const int j = sl[q].i1;
                            const float D = canon_dist<float>(xrow, Cs + j * CST, d);
                            if (D < bd || (D == bd && j < lab)) {
                                bd = D;
                                lab = j;
                            }
                        }
                    }
                    if (lab >= k) lab = bi;
                } else {
                    need_scan = true;
                }
            }
            // Exact scans, one row at a time by the whole wave: lane l takes
            // centroids l, l + 64, ...; strict < in centroid order per lane, then
            // the smallest value with ties to the lowest index; a NaN distance to
            // centroid 0 keeps label 0 (the exact-form tile kernel's rule).
            uint64_t scan_rows = __ballot(need_scan);
            while (scan_rows) {
                const int src = __ffsll((unsigned long long)scan_rows) - 1;  // a lane with h = 0
                scan_rows &= scan_rows - 1;
                const float* fx = in.X + (t0 + wave * 32 + src) * int64_t(d);
                float bv = finf;
                int bj = 0x7fffffff;
                bool nan0 = false;
                for (int j = lane; j < k; j += 64) {
                    const float D = canon_dist<float>(fx, Cs + j * CST, d);
                    if (j == 0 && D != D) nan0 = true;
                    if (D < bv) {
                        bv = D;
                        bj = j;
                    }
                }
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    const float ob = __shfl_xor(bv, off);
                    const int oj = __shfl_xor(bj, off);
                    if (ob < bv || (ob == bv && oj < bj)) {
                        bv = ob;
                        bj = oj;
                    }
                }
                const bool n0 = __shfl(nan0 ? 1 : 0, 0) != 0;
                if (lane == src) lab = (n0 || bj == 0x7fffffff) ? 0 : bj;
            }
            if (h == 0 && ok) {
                if (to_user) {
                    in.labels_out[row] = lab;
                } else {
                    labels[in.row0 + row] = lab;
                    if (skeys) skeys[in.row0 + row] = (uint32_t(in.index) << lbits) | uint32_t(lab);
                }
            }
        }
    }
}

// Autotuner inputs: instance i's rows are its centroids (row % k) plus uniform
// noise in [-1.75, 1.75) per feature, so that rows lie near one centroid as in
// clustered data.
__global__ void batch_tune_points_kernel(float* __restrict__ X, const float* __restrict__ C, int ni,
                                         int64_t n, int d, int k, uint32_t seed) {
    const int64_t total = int64_t(ni) * n * d;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += int64_t(gridDim.x) * blockDim.x) {
        const int64_t r = e / d;
        const int f = int(e - r * d);
        const int64_t inst = r / n;
        const int64_t j = (r - inst * n) % k;
        uint32_t h = uint32_t(e) * 0x9E3779B1u ^ uint32_t(e >> 32) ^ seed;
        h ^= h >> 15;
        h *= 0x85EBCA77u;
        h ^= h >> 13;
        X[e] = C[(inst * k + j) * d + f] + (float(h >> 8) * (3.5f / 16777216.0f) - 1.75f);
    }
}

// Launch kinds of a plan: the VALU tiles 0..6, then the matrix-core kernel for
// DP = 32, 64, 128, 16 (7, 8, 9, 10); rows per work item of each.
constexpr int BATCH_LAUNCH_KINDS = 11;
__host__ __device__ constexpr int batch_kind_rows(int kind) {
    return kind < BATCH_TILE_MFMA ? batch_tile_rows(kind) : BATCH_MFMA_ROWS;
}

// synthetic finite floats in [-10, 10) from a hash of the index (autotuner inputs)
__global__ void batch_tune_fill_kernel(float* __restrict__ p, int64_t count, uint32_t seed) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += int64_t(gridDim.x) * blockDim.x) {
        uint32_t h = uint32_t(i) * 0x9E3779B1u ^ uint32_t(i >> 32) ^ seed;
        h ^= h >> 15;
        h *= 0x85EBCA77u;
        h ^= h >> 13;
        p[i] = float(h >> 8) * (20.0f / 16777216.0f) - 10.0f;
    }
}

// label histogram: BATCH_CNT_ROWS rows of one instance per workgroup, added to
// counts (when not null) and written per block to bhist (when not null)
__global__ __launch_bounds__(256) void batch_count_kernel(const batch_inst* __restrict__ I, int B,
                                                          const int64_t* __restrict__ pre,
                                                          const int32_t* __restrict__ labels,
                                                          int32_t* __restrict__ counts,
                                                          int32_t* __restrict__ bhist) {
    __shared__ int hist[BATCH_HIST_MAX];
    const int64_t total = pre[B];
    for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
        const int ii = batch_find(pre, B, b);
        const batch_inst in = I[ii];
        const int64_t r0 = (b - pre[ii]) * BATCH_CNT_ROWS;
        const int nr = int(min<int64_t>(BATCH_CNT_ROWS, in.n - r0));
        const bool lds = in.k <= BATCH_HIST_MAX;
        __syncthreads();
        if (lds)
            for (int j = threadIdx.x; j < in.k; j += 256) hist[j] = 0;
        __syncthreads();
        // every label of the thread loaded before the first atomic
        constexpr int RPT = BATCH_CNT_ROWS / 256;
        int lv[RPT];
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
            const int r = threadIdx.x + 256 * u;
            lv[u] = r < nr ? labels[in.row0 + r0 + r] : -1;
        }
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
            const int lab = lv[u];
            if (lab >= 0 && lab < in.k) {
                if (lds) atomicAdd(&hist[lab], 1);
                else atomicAdd(&counts[in.k0 + lab], 1);
            }
        }
        __syncthreads();
        if (lds)
            for (int j = threadIdx.x; j < in.k; j += 256) {
                if (counts && hist[j]) atomicAdd(&counts[in.k0 + j], hist[j]);
                // the block's histogram, for the counting sort (bhist: every k <= BATCH_HIST_MAX)
                if (bhist) bhist[in.h0 + (b - pre[ii]) * in.k + j] = hist[j];
            }
    }
}

// ---------------------------------------------------------------------------
// Stable counting sort of the row ids by label within each instance (the order
// rocPRIM's radix sort of (instance, label) keys gives: rows of a cluster in
// ascending row order), for batches whose k are all at most BATCH_HIST_MAX.
// 1. batch_count_kernel writes each histogram block's label counts (bhist).
// 2. batch_scan_kernel sums them into the counts and, after the scan, turns them
//    into each block's first destination per label: the label's member offset
//    plus the counts of the instance's earlier blocks.
// 3. batch_scatter_kernel: one workgroup per block.  Each wave ranks the rows of
//    its quarter of the block within their label (one ballot per label bit per
//    64 rows, in row order; a per-wave LDS counter per label carries the runs
//    from one group of 64 rows to the next), the waves' counts give every row a
//    slot in the block's label-ordered LDS staging array, and the staging array
//    is written out in slot order, so each label's rows of the block leave as
//    one contiguous run.  Labels outside [0, k) are skipped, as the histogram
//    skips them.
// ---------------------------------------------------------------------------
// One large instance (the single-instance engine): its block destinations in
// two parallel steps instead of batch_scan_kernel's per-label walk over every
// block.  The nb blocks form G groups of GB consecutive blocks.
// sort_group_sums_kernel: gsum[g k + j] = label j's count over group g's blocks.
// sort_block_bases_kernel: thread (g, j) starts at offsets[j] plus label j's
// counts in the earlier groups and turns its blocks' counts into destinations,
// in block order.  The result equals batch_scan_kernel's.
__global__ __launch_bounds__(256) void sort_group_sums_kernel(const int32_t* __restrict__ bhist, int64_t nb,
                                                              int k, int G, int GB,
                                                              int32_t* __restrict__ gsum) {
    const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (e >= int64_t(G) * k) return;
    const int g = int(e / k), j = int(e - int64_t(g) * k);
    const int64_t b0 = int64_t(g) * GB, b1 = min<int64_t>(nb, b0 + GB);
    int32_t s = 0;
    int64_t lb = b0;
    for (; lb + 4 <= b1; lb += 4) {
        const int32_t h0 = bhist[lb * k + j], h1 = bhist[(lb + 1) * k + j], h2 = bhist[(lb + 2) * k + j],
                      h3 = bhist[(lb + 3) * k + j];
        s += h0 + h1 + h2 + h3;
    }
    for (; lb < b1; ++lb) s += bhist[lb * k + j];
    gsum[e] = s;
}

// counts[j] = label j's count over every block (the column sums of gsum), in
// place of one global atomic per (block, label)
__global__ __launch_bounds__(256) void sort_counts_kernel(const int32_t* __restrict__ gsum, int k, int G,
                                                          int32_t* __restrict__ counts) {
    const int j = int(blockIdx.x) * 256 + int(threadIdx.x);
    if (j >= k) return;
    int32_t s = 0;
#pragma unroll 16
    for (int g = 0; g < G; ++g) s += gsum[int64_t(g) * k + j];
    counts[j] = s;
}

__global__ __launch_bounds__(256) void sort_block_bases_kernel(int32_t* __restrict__ bhist, int64_t nb, int k,
                                                               int G, int GB, const int32_t* __restrict__ gsum,
                                                               const int64_t* __restrict__ offsets) {
    const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (e >= int64_t(G) * k) return;
    const int g = int(e / k), j = int(e - int64_t(g) * k);
    int32_t base = int32_t(offsets[j]);
#pragma unroll 16
    for (int q = 0; q < g; ++q) base += gsum[int64_t(q) * k + j];
    const int64_t b0 = int64_t(g) * GB, b1 = min<int64_t>(nb, b0 + GB);
    // eight blocks' counts loaded before their destinations are written
    int64_t lb = b0;
    for (; lb + 8 <= b1; lb += 8) {
        int32_t c[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) c[u] = bhist[(lb + u) * k + j];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            bhist[(lb + u) * k + j] = base;
            base += c[u];
        }
    }
    for (; lb < b1; ++lb) {
        const int32_t c = bhist[lb * k + j];
        bhist[lb * k + j] = base;
        base += c;
    }
}

// LDS bytes of batch_scatter_kernel for labels below kcap
__host__ __device__ constexpr size_t batch_scatter_lds_bytes(int kcap) {
    return (size_t(6) * kcap + BATCH_CNT_ROWS) * sizeof(int32_t);
}

__global__ __launch_bounds__(256) void batch_scatter_kernel(const batch_inst* __restrict__ I, int B,
                                                            const int64_t* __restrict__ pre,
                                                            const int32_t* __restrict__ labels,
                                                            const int32_t* __restrict__ bhist,
                                                            int32_t* __restrict__ perm, int kcap) {
    // LDS: wcnt[4][kcap] (each wave's label counts, then its first local slot per
    // label), ls[kcap] (label's first local slot), gb[kcap] (label's destination),
    // stage[BATCH_CNT_ROWS] (local slot -> label << 16 | row in block)
    extern __shared__ int32_t slds[];
    int32_t* ls = slds + 4 * kcap;
    int32_t* gb = ls + kcap;
    int32_t* stage = gb + kcap;
    __shared__ int32_t wsum[4];
    constexpr int G = BATCH_CNT_ROWS / 256;  // groups of 64 rows per wave
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    int32_t* wc = slds + wave * kcap;
    const uint64_t below = (uint64_t(1) << lane) - 1u;
    const int64_t total = pre[B];
    for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
        const int ii = batch_find(pre, B, b);
        const batch_inst in = I[ii];
        const int k = in.k;
        const int64_t lb = b - pre[ii];
        const int64_t r0 = lb * BATCH_CNT_ROWS;
        const int nr = int(min<int64_t>(BATCH_CNT_ROWS, in.n - r0));
        int lbits = 0;
        while ((1 << lbits) < k) ++lbits;
        const int32_t* lab_in = labels + in.row0 + r0;
        const int wr0 = wave * (BATCH_CNT_ROWS / 4);  // the wave's quarter of the block
        int32_t lab[G], rk[G];
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const int r = wr0 + q * 64 + lane;
            lab[q] = r < nr ? lab_in[r] : -1;
        }
        for (int j = lane; j < k; j += 64) wc[j] = 0;
        // 1. rank of each row among the earlier rows of the wave's quarter with its label
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const bool ok = lab[q] >= 0 && lab[q] < k;
            uint64_t same = __ballot(ok);
            for (int bit = 0; bit < lbits; ++bit) {
                const bool on = ok && ((lab[q] >> bit) & 1);
                const uint64_t m = __ballot(on);
                same &= on ? m : ~m;
            }
            const int rank = __popcll(same & below);
            rk[q] = -1;
            if (ok) {
                rk[q] = wc[lab[q]] + rank;
                if (rank == 0) wc[lab[q]] = rk[q] + __popcll(same);
            }
        }
        __syncthreads();
        // 2. local slots: labels in order, within a label the waves in order
        const int m = (k + 255) / 256;
        const int j0 = min(k, t * m), j1 = min(k, j0 + m);
        int sum = 0;
        for (int j = j0; j < j1; ++j) sum += slds[j] + slds[kcap + j] + slds[2 * kcap + j] + slds[3 * kcap + j];
        int incl = sum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        int excl = incl - sum;
        for (int w = 0; w < wave; ++w) excl += wsum[w];
        const int nvalid = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        const int32_t* base = bhist + in.h0 + lb * k;
        for (int j = j0; j < j1; ++j) {
            ls[j] = excl;
            gb[j] = base[j];
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const int c = slds[w * kcap + j];
                slds[w * kcap + j] = excl;
                excl += c;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < G; ++q)
            if (rk[q] >= 0) stage[wc[lab[q]] + rk[q]] = (lab[q] << 16) | (wr0 + q * 64 + lane);
        __syncthreads();
        // 3. the block's rows of each label to their destination run, slot order
        int32_t* out = perm + in.row0;
        const int32_t id0 = int32_t(in.row0 + r0);
        for (int p = t; p < nvalid; p += 256) {
            const int v = stage[p];
            const int l = v >> 16;
            out[gb[l] + (p - ls[l])] = id0 + (v & 0xffff);
        }
        __syncthreads();
    }
}

// Per instance (one workgroup each): member offsets, chunk bases and the list of
// empty clusters from the counts, as scan_clusters_kernel; meta[2i] = #empty,
// meta[2i+1] = #chunks.  With bhist (the counting sort), the counts are first
// summed from the per-block histograms and written to counts, and after the scan
// each label's per-block counts become that block's first destination (the
// label's member offset plus the counts of the instance's earlier blocks).
// The last workgroup to finish (a ticket counter, reset by that workgroup) writes
// the exclusive prefix over instances of their chunk counts: qpre[i] = first
// global chunk of instance i, qpre[B] = total; and, when dpre is not null, the
// same prefix of the d64 blocks (BATCH_D64_ROWS rows) of the instances that have
// empty clusters (0 blocks for the others), so the d64 pass visits only those.
__global__ __launch_bounds__(1024) void batch_scan_kernel(const batch_inst* __restrict__ I, int B,
                                                          int32_t* __restrict__ counts,
                                                          int32_t* __restrict__ bhist,
                                                          int64_t* __restrict__ offsets,
                                                          int32_t* __restrict__ chunk_base,
                                                          int32_t* __restrict__ empty_list,
                                                          int32_t* __restrict__ meta,
                                                          int64_t* __restrict__ qpre,
                                                          int64_t* __restrict__ dpre,
                                                          uint32_t* __restrict__ ticket) {
    __shared__ int64_t s_off[1024];
    __shared__ int32_t s_chk[1024];
    __shared__ int32_t s_emp[1024];
    __shared__ int64_t carry_off;
    __shared__ int32_t carry_chk, carry_emp;
    __shared__ bool s_last;
    const int t = threadIdx.x;
    for (int ii = blockIdx.x; ii < B; ii += gridDim.x) {
        const batch_inst in = I[ii];
        const int k = in.k;
        int32_t* cnt = counts + in.k0;
        int64_t* off = offsets + in.k0 + ii;
        int32_t* cb = chunk_base + in.k0 + ii;
        int32_t* emp = empty_list + in.k0;
        const int64_t nb = (in.n + BATCH_CNT_ROWS - 1) / BATCH_CNT_ROWS;
        __syncthreads();
        if (t == 0) {
            carry_off = 0;
            carry_chk = 0;
            carry_emp = 0;
        }
        __syncthreads();
        for (int base = 0; base < k; base += 1024) {
            const int j = base + t;
            int32_t c = 0;
            if (j < k) {
                if (bhist) {
                    const int32_t* h = bhist + in.h0 + j;
                    int64_t lb = 0;
                    for (; lb + 4 <= nb; lb += 4) {
                        const int32_t h0 = h[lb * k], h1 = h[(lb + 1) * k], h2 = h[(lb + 2) * k],
                                      h3 = h[(lb + 3) * k];
                        c += h0 + h1 + h2 + h3;
                    }
                    for (; lb < nb; ++lb) c += h[lb * k];
                    cnt[j] = c;
                } else {
                    c = cnt[j];
                }
            }
            s_off[t] = c;
            s_chk[t] = j < k ? (c + REDUCE_CHUNK - 1) / REDUCE_CHUNK : 0;
            s_emp[t] = (j < k && c == 0) ? 1 : 0;
            __syncthreads();
            for (int o = 1; o < 1024; o <<= 1) {
                int64_t a = 0;
                int32_t bch = 0, e = 0;
                if (t >= o) {
                    a = s_off[t - o];
                    bch = s_chk[t - o];
                    e = s_emp[t - o];
                }
                __syncthreads();
                s_off[t] += a;
                s_chk[t] += bch;
                s_emp[t] += e;
                __syncthreads();
            }
            if (j < k) {
                const int64_t first = carry_off + s_off[t] - c;
                off[j] = first;
                cb[j] = carry_chk + s_chk[t] - (c + REDUCE_CHUNK - 1) / REDUCE_CHUNK;
                if (c == 0) emp[carry_emp + s_emp[t] - 1] = j;
                if (bhist) {
                    int32_t* h = bhist + in.h0 + j;
                    int32_t dst = int32_t(first);
                    for (int64_t lb = 0; lb < nb; ++lb) {
                        const int32_t hc = h[lb * k];
                        h[lb * k] = dst;
                        dst += hc;
                    }
                }
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
            off[k] = carry_off;
            cb[k] = carry_chk;
            meta[2 * ii] = carry_emp;
            meta[2 * ii + 1] = carry_chk;
        }
    }
    // the last workgroup to finish: chunk prefix over the instances
    __syncthreads();
    if (t == 0) {
        __threadfence();  // this workgroup's meta before its ticket
        s_last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();  // every workgroup's meta after the last ticket
    for (int which = 0; which < 2; ++which) {
        int64_t* out = which == 0 ? qpre : dpre;
        if (!out) continue;
        int64_t carry = 0;
        for (int base = 0; base < B; base += 1024) {
            const int i = base + t;
            int64_t v = 0;
            if (i < B) {
                if (which == 0)
                    v = meta[2 * i + 1];
                else if (meta[2 * i] > 0)
                    v = (I[i].n + BATCH_D64_ROWS - 1) / BATCH_D64_ROWS;
            }
            s_off[t] = v;
            __syncthreads();
            for (int o = 1; o < 1024; o <<= 1) {
                const int64_t a = t >= o ? s_off[t - o] : 0;
                __syncthreads();
                s_off[t] += a;
                __syncthreads();
            }
            if (i < B) out[i] = carry + s_off[t] - v;
            carry += s_off[1023];
            __syncthreads();
        }
        if (t == 0) out[B] = carry;
    }
    if (t == 0) *ticket = 0u;  // ready for the next launch
}

// Chunk sums of the instances whose rows allow 16-byte reads (d % 4 == 0,
// d <= 256, X 16-byte aligned).  A wave packs S = 64 / lpc consecutive chunks
// of one instance into sub-waves of lpc lanes; within a sub-wave lane (g, f4)
// owns features 4 f4 .. 4 f4 + 3 of group g's members (d <= 32: the
// reduce_groups(d) interleaved partials of the scalar kernel, G * d / 4 <= 16
// lanes) or of every member (d > 32: d / 4 lanes), so each feature is summed over
// the same members in the same order as batch_reduce_kernel's and the partials
// are identical.  Waves take runs of consecutive chunks (up to RUN), round
// robin; member ids are loaded one step ahead of the row reads.
__global__ __launch_bounds__(256) void batch_reduce_vec_kernel(const batch_inst* __restrict__ I, int B,
                                                               const int64_t* __restrict__ qpre,
                                                               const int32_t* __restrict__ perm,
                                                               const int64_t* __restrict__ offsets,
                                                               const int32_t* __restrict__ chunk_base,
                                                               double* __restrict__ part) {
    typedef float F4 __attribute__((ext_vector_type(4)));
    constexpr int RUN = 16;
    constexpr int U = 8;
    const int lane = threadIdx.x & 63;
    const int64_t gw = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = (int64_t(gridDim.x) * blockDim.x) >> 6;
    const int64_t total = qpre[B];
    // runs of RUN chunks when there is work for every wave, shorter runs (down to
    // one chunk per wave) when there is not
    const int64_t run = max<int64_t>(1, min<int64_t>(RUN, total / nw));
    for (int64_t q0 = gw * run; q0 < total; q0 += nw * run) {
        const int64_t q1 = min<int64_t>(q0 + run, total);
        int ii = batch_find(qpre, B, q0);
        int64_t q = q0;
        while (q < q1) {
            while (q >= qpre[ii + 1]) ++ii;  // every instance has chunks
            const batch_inst in = I[ii];
            const int64_t inst_end = min<int64_t>(q1, qpre[ii + 1]);
            const int d = in.d;
            if (!(in.vec && d <= 256)) {  // batch_reduce_kernel's chunks
                q = inst_end;
                continue;
            }
            const int d4 = d >> 2;
            const int lpc = d <= 64 ? 16 : (d <= 128 ? 32 : 64);  // lanes per chunk
            const int S = 64 / lpc;
            const int64_t qe = min<int64_t>(inst_end, q + S);
            const int sub = lane / lpc, sl = lane - sub * lpc;
            const int64_t myq = q + sub;
            const bool has = myq < qe;
            const int G = reduce_groups(d);
            const int g = d <= 32 ? sl / d4 : 0, f4 = d <= 32 ? sl - g * d4 : sl;
            const bool active = has && (d <= 32 ? g < G : sl < d4);
            const int Sm = d <= 32 ? G : 1;  // member stride of a lane
            int64_t m_begin = 0, m_end = 0;
            double* out = nullptr;
            if (has) {
                const int32_t lq = int32_t(myq - qpre[ii]);
                const int64_t* off = offsets + in.k0 + ii;
                const int32_t* cb = chunk_base + in.k0 + ii;
                int lo = 0, hi = in.k;  // cb[k] = chunks > lq
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (cb[mid] <= lq) lo = mid; else hi = mid;
                }
                m_begin = off[lo] + int64_t(lq - cb[lo]) * REDUCE_CHUNK;
                m_end = min(m_begin + REDUCE_CHUNK, off[lo + 1]);
                out = part + in.p0 + int64_t(lq) * d;
            }
            const int32_t* pm = perm + in.row0;
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
            if (active) {
                int64_t m = m_begin;
                int32_t r[U];
                if (m + U * Sm <= m_end) {
#pragma unroll
                    for (int u = 0; u < U; ++u) r[u] = pm[m + u * Sm + g];
                }
                for (; m + U * Sm <= m_end; m += U * Sm) {
                    const bool more = m + 2 * U * Sm <= m_end;
                    int32_t rn[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) rn[u] = more ? pm[m + U * Sm + u * Sm + g] : 0;
                    F4 xv[U];
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        xv[u] = *reinterpret_cast<const F4*>(in.X + (int64_t(r[u]) - in.row0) * d + 4 * f4);
#pragma unroll
                    for (int u = 0; u < U; ++u)
#pragma unroll
                        for (int e = 0; e < 4; ++e) acc[e] = acc[e] + double(xv[u][e]);
#pragma unroll
                    for (int u = 0; u < U; ++u) r[u] = rn[u];
                }
                for (; m < m_end; m += Sm) {
                    const int64_t mm = m + g;
                    if (mm < m_end) {
                        const int64_t row = int64_t(pm[mm]) - in.row0;
                        const F4 x4 = *reinterpret_cast<const F4*>(in.X + row * d + 4 * f4);
#pragma unroll
                        for (int e = 0; e < 4; ++e) acc[e] = acc[e] + double(x4[e]);
                    }
                }
            }
            // every lane reaches the shuffles: group partials into group 0, in
            // group order, within the sub-wave
            double tot[4] = {acc[0], acc[1], acc[2], acc[3]};
            if (d <= 32) {
                for (int gg = 1; gg < G; ++gg)
#pragma unroll
                    for (int e = 0; e < 4; ++e) tot[e] = tot[e] + __shfl(acc[e], sub * lpc + gg * d4 + f4);
            }
            if (active && g == 0)
#pragma unroll
                for (int e = 0; e < 4; ++e) out[4 * f4 + e] = tot[e];
            q = qe;
        }
    }
}

// Chunk sums, one wave per chunk of REDUCE_CHUNK members of one cluster of one
// instance, in the order of reduce_chunks_kernel: d <= 32 uses reduce_groups(d)
// interleaved partials added in group order; otherwise each feature is summed
// over the members in order (64 features per pass).  perm holds global row ids
// (row0 + local row) sorted by label within each instance.
__global__ __launch_bounds__(256) void batch_reduce_kernel(const batch_inst* __restrict__ I, int B,
                                                           const int64_t* __restrict__ qpre,
                                                           const int32_t* __restrict__ perm,
                                                           const int64_t* __restrict__ offsets,
                                                           const int32_t* __restrict__ chunk_base,
                                                           double* __restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int64_t gw = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = (int64_t(gridDim.x) * blockDim.x) >> 6;
    const int64_t total = qpre[B];
    // waves take runs of consecutive chunks, round robin: the owning instance
    // and cluster are searched once per run and then advanced chunk by chunk;
    // runs are RUN chunks long when there is work for every wave, shorter when not
    constexpr int RUN = 16;
    const int64_t run = max<int64_t>(1, min<int64_t>(RUN, total / nw));
    for (int64_t q0 = gw * run; q0 < total; q0 += nw * run) {
        const int64_t q1 = min<int64_t>(q0 + run, total);
        int ii = batch_find(qpre, B, q0);
        int j = -1;
        for (int64_t q = q0; q < q1; ++q) {
            while (q >= qpre[ii + 1]) {  // next instance (every instance has chunks)
                ++ii;
                j = -1;
            }
            const batch_inst in = I[ii];
            const int d = in.d, k = in.k;
            if (in.vec && d <= 256) {
                // batch_reduce_vec_kernel sums this instance's chunks
                q = min<int64_t>(q1, qpre[ii + 1]) - 1;
                continue;
            }
            const int32_t lq = int32_t(q - qpre[ii]);
            const int64_t* off = offsets + in.k0 + ii;
            const int32_t* cb = chunk_base + in.k0 + ii;
            if (j < 0) {
                int lo = 0, hi = k;  // cb[k] = chunks > lq
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (cb[mid] <= lq) lo = mid; else hi = mid;
                }
                j = lo;
            }
            while (cb[j + 1] <= lq) ++j;  // clusters without chunks are passed over
            const int64_t m_begin = off[j] + int64_t(lq - cb[j]) * REDUCE_CHUNK;
            const int64_t m_end = min(m_begin + REDUCE_CHUNK, off[j + 1]);
            const int32_t* pm = perm + in.row0;
            double* out = part + in.p0 + int64_t(lq) * d;
            if (d <= 32) {
                const int G = reduce_groups(d);
                const int g = lane / d, f0 = lane % d;
                const bool active = g < G;
                double acc = 0.0;
                int64_t m = m_begin;
                // eight steps of G members: ids and values loaded first, added in order
                for (; m + 8 * G <= m_end; m += 8 * G) {
                    if (active) {
                        int64_t r[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) r[u] = int64_t(pm[m + u * G + g]) - in.row0;
                        double v[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) v[u] = double(in.X[r[u] * d + f0]);
#pragma unroll
                        for (int u = 0; u < 8; ++u) acc = acc + v[u];
                    }
                }
                for (; m < m_end; m += G) {
                    const int64_t mm = m + g;
                    if (active && mm < m_end) {
                        const int64_t row = int64_t(pm[mm]) - in.row0;
                        acc = acc + double(in.X[row * d + f0]);
                    }
                }
                double tot = acc;
                for (int gg = 1; gg < G; ++gg) {
                    const double o = __shfl(acc, gg * d + f0);
                    tot = tot + o;
                }
                if (g == 0) out[f0] = tot;
            } else {
                for (int fc = 0; fc < d; fc += 64) {
                    const int f = fc + lane;
                    double acc = 0.0;
                    int64_t m = m_begin;
                    for (; m + 16 <= m_end; m += 16) {
                        int64_t r[16];
#pragma unroll
                        for (int u = 0; u < 16; ++u) r[u] = int64_t(pm[m + u]) - in.row0;
                        if (f < d) {
                            double v[16];
#pragma unroll
                            for (int u = 0; u < 16; ++u) v[u] = double(in.X[r[u] * d + f]);
#pragma unroll
                            for (int u = 0; u < 16; ++u) acc = acc + v[u];
                        }
                    }
                    for (; m < m_end; ++m) {
                        const int64_t r1 = int64_t(pm[m]) - in.row0;
                        if (f < d) acc = acc + double(in.X[r1 * d + f]);
                    }
                    if (f < d) out[f] = acc;
                }
            }
        }
    }
}

// C_next[j][f] = (float)(sum of the cluster's chunk partials in order / count);
// empty clusters are left for batch_fill_empty_kernel.  One thread per centroid
// element (prefix table over k * d).
__global__ __launch_bounds__(256) void batch_finalize_kernel(const batch_inst* __restrict__ I, int B,
                                                             const int64_t* __restrict__ pre,
                                                             const int32_t* __restrict__ counts,
                                                             const int32_t* __restrict__ chunk_base,
                                                             const double* __restrict__ part,
                                                             float* __restrict__ Cnext) {
    const int64_t total = pre[B];
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += int64_t(gridDim.x) * blockDim.x) {
        const int ii = batch_find(pre, B, e);
        const batch_inst in = I[ii];
        const int64_t le = e - pre[ii];
        const int j = int(le / in.d), f = int(le - int64_t(j) * in.d);
        const int32_t cnt = counts[in.k0 + j];
        if (cnt == 0) continue;
        const int32_t* cb = chunk_base + in.k0 + ii;
        const double* pp = part + in.p0;
        double s = 0.0;
        int q = cb[j];
        const int q1 = cb[j + 1];
        for (; q + 8 <= q1; q += 8) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = pp[int64_t(q + u) * in.d + f];
#pragma unroll
            for (int u = 0; u < 8; ++u) s = s + v[u];
        }
        for (; q < q1; ++q) s = s + pp[int64_t(q) * in.d + f];
        Cnext[in.c0 + le] = float(s / double(cnt));
    }
}

// FP64 squared distance of one row x to its labelled centroid c: sum over the
// features in order of (double(x) - double(c))^2, each sub, mul and add rounded
// separately (the single-instance d64 and objective order).  The thread reads the
// row and the centroid directly, a 128-byte line per step with 16-byte loads when
// vec (d % 4 == 0, 16-byte aligned rows and centroids).
__device__ __forceinline__ double batch_row_d64(const float* __restrict__ x, const float* __restrict__ c,
                                                int d, bool vec) {
    typedef float F4 __attribute__((ext_vector_type(4)));
    double acc = 0.0;
    int f = 0;
    if (vec) {
        for (; f + 32 <= d; f += 32) {
            F4 xv[8], cv[8];
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                xv[v] = *reinterpret_cast<const F4*>(x + f + 4 * v);
                cv[v] = *reinterpret_cast<const F4*>(c + f + 4 * v);
            }
#pragma unroll
            for (int v = 0; v < 8; ++v)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const double diff = double(xv[v][e]) - double(cv[v][e]);
                    const double sq = diff * diff;
                    acc = acc + sq;
                }
        }
        for (; f + 4 <= d; f += 4) {
            const F4 x4 = *reinterpret_cast<const F4*>(x + f);
            const F4 c4 = *reinterpret_cast<const F4*>(c + f);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const double diff = double(x4[e]) - double(c4[e]);
                const double sq = diff * diff;
                acc = acc + sq;
            }
        }
    }
    for (; f < d; ++f) {
        const double diff = double(x[f]) - double(c[f]);
        const double sq = diff * diff;
        acc = acc + sq;
    }
    return acc;
}

// d64 of every row of the instances that have empty clusters (meta[2i] > 0), or
// of every instance when meta is null; pre: the instances' d64 block prefix (with
// meta, the scan's prefix over the instances with empty clusters, so the others
// take no blocks); labels: concatenated (row0), or each instance's labels_out
// when labels is null; BATCH_D64_ROWS rows per workgroup
__global__ __launch_bounds__(256) void batch_d64_kernel(const batch_inst* __restrict__ I, int B,
                                                        const int64_t* __restrict__ pre,
                                                        const int32_t* __restrict__ meta,
                                                        const float* __restrict__ C,
                                                        const int32_t* __restrict__ labels,
                                                        double* __restrict__ d64) {
    static_assert(BATCH_D64_ROWS == 256, "one row per thread");
    const int64_t total = pre[B];
    for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
        const int ii = batch_find(pre, B, b);
        if (meta && meta[2 * ii] == 0) continue;  // uniform over the workgroup
        const batch_inst in = I[ii];
        const int64_t r0 = (b - pre[ii]) * BATCH_D64_ROWS;
        const int nr = int(min<int64_t>(BATCH_D64_ROWS, in.n - r0));
        const int32_t* lab = labels ? labels + in.row0 : in.labels_out;
        const int t = threadIdx.x;
        if (t < nr) {
            const int64_t r = r0 + t;
            int j = lab[r];
            j = (j >= 0 && j < in.k) ? j : 0;
            // centroid rows are 16-byte aligned when d % 4 == 0 (c0 is a multiple of 4)
            d64[in.row0 + r] = batch_row_d64(in.X + r * in.d, C + in.c0 + int64_t(j) * in.d, in.d, in.vec != 0);
        }
    }
}

// Empty clusters of one instance per workgroup: pick e = 0 .. m-1 takes the row
// with the largest d64 (ties: lowest row) after pick e-1 in that order, and the
// e-th empty cluster (ascending index) receives that row -- the policy of
// empty_cluster_handling_dpc.cpp:26-114 with its tie order pinned.  An empty
// cluster left without a row (every remaining d64 is NaN) keeps its centroid.
__global__ __launch_bounds__(256) void batch_fill_empty_kernel(const batch_inst* __restrict__ I, int B,
                                                               const int32_t* __restrict__ meta,
                                                               const int32_t* __restrict__ empty_list,
                                                               const double* __restrict__ d64,
                                                               const float* __restrict__ Ccur,
                                                               float* __restrict__ Cnext) {
    __shared__ double sd[256];
    __shared__ int64_t sr[256];
    const int t = threadIdx.x;
    for (int ii = blockIdx.x; ii < B; ii += gridDim.x) {
        const int m = meta[2 * ii];
        if (m == 0) continue;
        const batch_inst in = I[ii];
        const double* dd = d64 + in.row0;
        double pd = __builtin_inf();
        int64_t pr = -1;
        for (int e = 0; e < m; ++e) {
            double bd = -1.0;
            int64_t br = -1;
            for (int64_t r = t; r < in.n; r += 256) {
                const double v = dd[r];
                const bool after = v < pd || (v == pd && r > pr);
                if (after && (br < 0 || v > bd)) {  // rows ascend per thread: first max wins
                    bd = v;
                    br = r;
                }
            }
            sd[t] = bd;
            sr[t] = br;
            __syncthreads();
            for (int s = 128; s > 0; s >>= 1) {
                if (t < s) {
                    const double od = sd[t + s];
                    const int64_t orr = sr[t + s];
                    if (orr >= 0 && (sr[t] < 0 || od > sd[t] || (od == sd[t] && orr < sr[t]))) {
                        sd[t] = od;
                        sr[t] = orr;
                    }
                }
                __syncthreads();
            }
            const double wd = sd[0];
            const int64_t wr = sr[0];
            __syncthreads();
            const int j = empty_list[in.k0 + e];
            for (int f = t; f < in.d; f += 256)
                Cnext[in.c0 + int64_t(j) * in.d + f] =
                    wr >= 0 ? in.X[wr * in.d + f] : Ccur[in.c0 + int64_t(j) * in.d + f];
            pd = wd;
            pr = wr;
        }
    }
}

// Objective chunk sums from the rows' d64 (batch_d64_kernel): OBJ_CHUNK rows of
// one instance per workgroup, in the order of objective_chunks_kernel (thread t
// sums rows t, t+256, ...; tree of the 256 partials).
__global__ __launch_bounds__(256) void batch_objective_kernel(const batch_inst* __restrict__ I, int B,
                                                              const int64_t* __restrict__ pre,
                                                              const double* __restrict__ d64,
                                                              double* __restrict__ chunk_sums) {
    __shared__ double part[256];
    const int t = threadIdx.x;
    const int64_t total = pre[B];
    for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
        const int ii = batch_find(pre, B, b);
        const batch_inst in = I[ii];
        const int64_t lq = b - pre[ii];
        const int64_t r0 = lq * OBJ_CHUNK;
        const int rows = int(min<int64_t>(OBJ_CHUNK, in.n - r0));
        const double* dv = d64 + in.row0 + r0;
        double p = 0.0;
        for (int s0 = 0; s0 < rows; s0 += 256)
            if (s0 + t < rows) p = p + dv[s0 + t];
        __syncthreads();
        part[t] = p;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (t < s) part[t] = part[t] + part[t + s];
            __syncthreads();
        }
        if (t == 0) chunk_sums[in.o0 + lq] = part[0];
    }
}

// per instance: objective = chunk sums added in chunk order (sum_sequential_kernel)
__global__ void batch_objective_sum_kernel(const batch_inst* __restrict__ I, int B,
                                           const double* __restrict__ chunk_sums,
                                           double* __restrict__ out) {
    for (int ii = blockIdx.x * blockDim.x + threadIdx.x; ii < B; ii += gridDim.x * blockDim.x) {
        const batch_inst in = I[ii];
        const int64_t count = (in.n + OBJ_CHUNK - 1) / OBJ_CHUNK;
        const double* v = chunk_sums + in.o0;
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
        out[ii] = s;
    }
}

// centroid copies over the k * d elements of every instance (prefix table):
// dir 0: C0 -> C (start), dir 1: C -> C_out (result; instances without C_out are
// skipped)
__global__ __launch_bounds__(256) void batch_centroid_copy_kernel(const batch_inst* __restrict__ I,
                                                                  int B,
                                                                  const int64_t* __restrict__ pre,
                                                                  float* __restrict__ C, int dir) {
    const int64_t total = pre[B];
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += int64_t(gridDim.x) * blockDim.x) {
        const int ii = batch_find(pre, B, e);
        const batch_inst in = I[ii];
        const int64_t le = e - pre[ii];
        if (dir == 0)
            C[in.c0 + le] = in.C0[le];
        else if (in.C_out != nullptr)
            in.C_out[le] = C[in.c0 + le];
    }
}

template <typename T>
__global__ void copy_kernel(const T* __restrict__ src, T* __restrict__ dst, int64_t count) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += int64_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

}  // namespace km
