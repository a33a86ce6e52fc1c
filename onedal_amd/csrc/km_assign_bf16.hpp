// km_assign_bf16.hpp -- BF16 matrix-core assignment with exact re-score.
//
// Distances for all n x k pairs are formed on v_mfma_f32_32x32x16_bf16 from a
// resident BF16 copy of X and BF16 centroids, FP32 accumulation.  The value a
// lane tracks is
//     v_j = fl(cnB_j - 2 * xb . cb_j),  cnB_j = fl(||c_j||^2 + B)
// (the MFMA accumulator starts at cnB_j and the A operand holds bf16(-2 c_j),
// an exact scaling).  B >= max ||x||^2 plus the error bound keeps every v_j
// positive, so the float bits order like the values and a key
//     key = (bits(v_j) with the low IB bits replaced by (block, register))
// turns the running top-2 into three integer ops per distance
// (and_or, med3, min; min3 handles two at a time).  Replacing IB low bits
// truncates v_j by less than 2^(IB-23) v_j, which the bound includes.
//
// Error bound (per row; X = ||x||, Xh = ||bf16(x)||, ex = ||x - bf16(x)||,
// Ch = max ||bf16(c_j)||, Ec = max ||c_j - bf16(c_j)||, C = max ||c_j||):
//   |x.c - xb.cb| <= Xh Ec + ex Ch + ex Ec
//   accumulation of d products and the start value: gamma_{d+1} (B + C^2 + 2 Xh Ch)
//   rounding of cnB: 2u (B + C^2); key truncation: 2^(IB-23) (B + C^2 + 2 Xh Ch)
// Certification and the candidate / collect / full re-score lists follow
// km_assign.hpp.  A row whose slots cannot certify its candidate set goes to
// the collect pass: the same BF16 distances are formed again for that row
// (gathered rows, same kernel) and every j with v_j <= T is recorded; the exact
// re-score then runs over those ids.
#pragma once

#include "km_assign.hpp"
#include "km_common.hpp"

namespace km {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int uintx4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even float -> bf16 bits (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
    const uint32_t b = __float_as_uint(f);
    return uint16_t((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// a + b rounded up, for a, b >= 0 (TwoSum gives the rounding error exactly)
__device__ __forceinline__ float add_round_up(float a, float b) {
    const float s = a + b;
    const float bb = s - a;
    const float err = (a - (s - bb)) + (b - bb);
    return err > 0.f ? __uint_as_float(__float_as_uint(s) + 1u) : s;
}

// per-iteration parameters of the BF16 pass, written on the device
struct bf16_params {
    double B;        // value shift
    double C;        // max ||c_j|| (up)
    double Ch;       // max ||bf16(c_j)|| (up)
    double Ec;       // max ||c_j - bf16(c_j)|| (up)
};

// Host-computed constants of the bound.
struct bf16_consts {
    double g_acc;    // gamma_{d+1} of the accumulation (doubled unit roundoff)
    double u;        // 2^-24
    double trunc;    // 2^(IB-23)
    double g_canon;  // g' of the exact-form distance
    int ib;          // index bits in the key
};

__device__ __forceinline__ double bf16_row_delta(const bf16_consts& k, const bf16_params& p,
                                                 double X, double ex) {
    const double Xh = X + ex;
    const double mag = p.B + p.C * p.C + 2.0 * Xh * p.Ch;
    const double es = Xh * p.Ec + ex * p.Ch + ex * p.Ec;
    return 2.0 * es + k.g_acc * mag + 2.0 * k.u * (p.B + p.C * p.C) + k.trunc * mag;
}

// threshold in the shifted value domain (see km_common.hpp band_threshold)
__device__ __forceinline__ double bf16_threshold(const bf16_consts& k, const bf16_params& p,
                                                 double X, double ex, double m1) {
    const double da = bf16_row_delta(k, p, X, ex);
    const double m1s = m1 - p.B;  // approximate ||c||^2 - 2 x.c of the best centroid
    const double dist_up = fmax(0.0, m1s + da + X * X);
    const double band = (2.0 * da + 2.0 * k.g_canon * dist_up) * (1.0 + 9.5367431640625e-07);
    return m1 + band + (fabs(m1) + X * X + p.C * p.C) * 4.44e-16;
}

// ---------------------------------------------------------------------------
// X -> padded BF16 rows (stride DPB), ||x|| and ||x - bf16(x)|| (rounded up),
// and their maxima.  One wave per row.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void convert_rows_bf16_kernel(
    const float* __restrict__ X, int64_t n, int d, int dpb, uint16_t* __restrict__ Xb,
    float* __restrict__ xnorm_up, float* __restrict__ enorm_up,
    unsigned long long* __restrict__ max_bits /* [0] max ||x||, [1] max ex */) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * int64_t(blockDim.x) + threadIdx.x) >> 6;
    const int64_t nwaves = (int64_t(gridDim.x) * blockDim.x) >> 6;
    double mx = 0.0, me = 0.0;
    for (int64_t row = wave; row < n; row += nwaves) {
        double s = 0.0, e = 0.0;
        for (int f = lane; f < dpb; f += 64) {
            const float v = f < d ? X[row * d + f] : 0.f;
            const uint16_t b = f2bf(v);
            Xb[row * dpb + f] = b;
            const double dv = double(v);
            const double de = dv - double(bf2f(b));
            s = fma(dv, dv, s);
            e = fma(de, de, e);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            s += __shfl_xor(s, off);
            e += __shfl_xor(e, off);
        }
        if (lane == 0) {
            const double r = sqrt(s) * (1.0 + 1e-12);
            const double re = sqrt(e) * (1.0 + 1e-12) + 1e-30;
            float fr = float(r), fe = float(re);
            if (double(fr) < r) fr = __uint_as_float(__float_as_uint(fr) + 1u);
            if (double(fe) < re) fe = __uint_as_float(__float_as_uint(fe) + 1u);
            xnorm_up[row] = fr;
            enorm_up[row] = fe;
            mx = fmax(mx, double(fr));
            me = fmax(me, double(fe));
        }
    }
    if (lane == 0) {
        atomicMax(&max_bits[0], (unsigned long long)__double_as_longlong(mx));
        atomicMax(&max_bits[1], (unsigned long long)__double_as_longlong(me));
    }
}

// ---------------------------------------------------------------------------
// centroids -> bf16(-2 c) padded rows, ||c||^2 (F), maxima of ||c||,
// ||bf16(c)||, ||c - bf16(c)||.  One wave per centroid.
// max_bits: [0] C, [1] Ch, [2] Ec
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void centroid_prep_bf16_kernel(
    const float* __restrict__ C, int k, int d, int dpb, uint16_t* __restrict__ Cb,
    float* __restrict__ cn, unsigned long long* __restrict__ max_bits) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    for (int j = wave; j < k; j += nwaves) {
        double s = 0.0, sh = 0.0, se = 0.0;
        for (int f = lane; f < dpb; f += 64) {
            const float v = f < d ? C[int64_t(j) * d + f] : 0.f;
            const uint16_t b = f2bf(v);
            Cb[int64_t(j) * dpb + f] = f2bf(-2.0f * bf2f(b));  // exact: -2 * bf16(c)
            const double dv = double(v), dh = double(bf2f(b)), de = dv - dh;
            s = fma(dv, dv, s);
            sh = fma(dh, dh, sh);
            se = fma(de, de, se);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            s += __shfl_xor(s, off);
            sh += __shfl_xor(sh, off);
            se += __shfl_xor(se, off);
        }
        if (lane == 0) {
            cn[j] = float(s);
            const double up = 1.0 + 1e-12;
            atomicMax(&max_bits[0], (unsigned long long)__double_as_longlong(sqrt(s) * up));
            atomicMax(&max_bits[1], (unsigned long long)__double_as_longlong(sqrt(sh) * up));
            atomicMax(&max_bits[2], (unsigned long long)__double_as_longlong(sqrt(se) * up + 1e-30));
        }
    }
}

// B and the bound parameters for this pass (one thread).
__global__ void bf16_params_kernel(const unsigned long long* __restrict__ cmax_bits,
                                   const unsigned long long* __restrict__ xmax_bits,
                                   bf16_consts kc, bf16_params* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    bf16_params p;
    p.C = __longlong_as_double((long long)cmax_bits[0]);
    p.Ch = __longlong_as_double((long long)cmax_bits[1]);
    p.Ec = __longlong_as_double((long long)cmax_bits[2]);
    const double X = __longlong_as_double((long long)xmax_bits[0]);
    const double ex = __longlong_as_double((long long)xmax_bits[1]);
    // B >= max ||x||^2 + the largest value error, so every v_j stays positive;
    // the error terms that scale with B are below 2^-12 B, hence the 1.01 factor
    p.B = 0.0;
    const double d0 = bf16_row_delta(kc, p, X, ex);
    const double B = 1.01 * (X * X + d0) + 2.0 * kc.trunc * (p.C * p.C + 2.0 * (X + ex) * p.Ch) + 1.0;
    // a float value, so the kernels add exactly this shift
    float fb = float(B);
    if (double(fb) < B) fb = __uint_as_float(__float_as_uint(fb) + 1u);
    p.B = double(fb);
    *out = p;
}

// cnB_j = ||c_j||^2 + B rounded up (v_j keeps its lower bound); padded ids get
// a huge start value so they are never chosen.  kp = 32 * ceil(k / 32) entries.
__global__ void cnb_kernel(const float* __restrict__ cn, int k, int kp,
                           const bf16_params* __restrict__ params_p, float* __restrict__ cnb) {
    const float Bf = float(params_p->B);
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < kp; j += gridDim.x * blockDim.x)
        cnb[j] = j < k ? add_round_up(cn[j], Bf) : __uint_as_float(0x7F000000u);
}

// Centroid blocks live in a ring of BF16_RING LDS buffers, filled BF16_RING-1
// blocks ahead with direct-to-LDS loads.
constexpr int BF16_RING = 4;

// 16 bytes per lane from global memory straight into LDS: lane l writes
// lds_wave_base + 16 l (wave-uniform base, through M0).  Issued as inline asm so
// the compiler does not insert its own conservative vmcnt(0) before every LDS
// read; ordering is by the explicit vmcnt waits and barriers of the caller.
__device__ __forceinline__ void lds_dma16(const void* gsrc, void* lds_wave_base) {
    // the base is wave-uniform; readfirstlane tells the compiler so (SGPR operand)
    const uint32_t m0 = __builtin_amdgcn_readfirstlane(uint32_t(reinterpret_cast<uintptr_t>(lds_wave_base)));
    asm volatile(
        "s_mov_b32 m0, %0\n\t"
        "s_nop 0\n\t"
        "global_load_lds_dwordx4 %1, off"
        :
        : "s"(m0), "v"(gsrc)
        : "memory", "m0");
}

// LDS bytes of assign_bf16_kernel: ring of 32 x dpb bf16 blocks + their cnB
__host__ inline size_t bf16_lds_bytes(int dpb) {
    return size_t(BF16_RING) * (size_t(32) * dpb * 2 + 32 * 4);
}

// ---------------------------------------------------------------------------
// Main / collect kernel.  Workgroup = 4 waves; each wave owns NB blocks of 32
// rows (B operand in VGPRs, 8 bf16 per step); centroid blocks of 32 rows (A
// operand) stream through an LDS ring by global_load_lds_dwordx4, three blocks
// ahead of use.  A block is stored unpadded, 16-byte chunk c of row r at chunk
// c ^ (r & SWZ), so the 32 lanes reading one chunk column hit distinct banks.
// Lane l = 32*h + r: half h covers features [h*DPB/2, (h+1)*DPB/2), 8 per step.
// Slots: lane keeps 4 (centroid block index mod 4); a row has 8.
// COLLECT: rows come from L.coll_rows; every j with v_j <= T is recorded.
// ---------------------------------------------------------------------------
template <int DPB, int NB, bool COLLECT>
__global__ __launch_bounds__(256, 2) void assign_bf16_kernel(
    const uint16_t* __restrict__ Xb, int64_t n, const uint16_t* __restrict__ Cb, int k,
    const float* __restrict__ cnb_g, const float* __restrict__ xnorm_up,
    const float* __restrict__ enorm_up, const bf16_params* __restrict__ params_p,
    bf16_consts kc, int32_t* __restrict__ labels, refine_lists L) {
    constexpr int HALF = DPB / 2;
    constexpr int STEPS = HALF / 8;
    constexpr int ROWS_PER_WAVE = 32 * NB;
    constexpr int ROWS_PER_WG = 4 * ROWS_PER_WAVE;
    constexpr int CPR = DPB / 8;                  // 16-byte chunks per centroid row
    constexpr int SWZ = (CPR < 16 ? CPR : 16) - 1;  // chunk swizzle mask
    constexpr int ROWB = DPB * 2;                 // bytes per centroid row
    constexpr int BLKB = 32 * ROWB;               // bytes per block
    constexpr int NCOPY = BLKB / 1024 + 1;        // 1 KB copies per block (+1: cnB)
    constexpr int CPW = (NCOPY + 3) / 4;          // copies per wave per block (uniform)
    constexpr int NSL = 4;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned char* ring = smem_raw;                                 // [RING][BLKB]
    unsigned char* ring_cn = smem_raw + size_t(BF16_RING) * BLKB;   // [RING][32] floats

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int r32 = lane & 31;
    const int h = lane >> 5;
    const bf16_params P = *params_p;
    const int mblocks = (k + 31) / 32;
    const uint32_t kmask = 0x7FFFFFFFu & ~((1u << kc.ib) - 1u);
    const int64_t nrows = COLLECT ? int64_t(L.counters[2]) : n;
    const int64_t ntiles = (nrows + ROWS_PER_WG - 1) / ROWS_PER_WG;

    // Direct-to-LDS copies of block mb into ring slot `slot`.  Copy q (q < NCOPY-1)
    // fills LDS bytes [q KB, (q+1) KB) of the block; lane l writes 16-byte chunk
    // p = 64 q + l, i.e. physical row p / CPR, physical chunk p % CPR, which holds
    // logical chunk (p % CPR) ^ (row & SWZ).  Copy NCOPY-1 moves the 32 cnB values.
    // Every wave issues exactly CPW copies per block (repeating copy 0 when it has
    // fewer; the repeat writes the same bytes), so vmcnt waits are uniform.
    auto issue_block = [&](int mb, int slot) {
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            int q = wave + 4 * c;
            if (q >= NCOPY) q = 0;
            if (q < NCOPY - 1) {
                const int p = q * 64 + lane;
                const int prow = p / CPR, pc = p % CPR;
                const int lc = pc ^ (prow & SWZ);
                const int cj = min(mb * 32 + prow, k - 1);
                const uint16_t* src = Cb + int64_t(cj) * DPB + lc * 8;
                lds_dma16(src, ring + slot * BLKB + q * 1024);
            } else if (lane < 8) {
                lds_dma16(cnb_g + mb * 32 + lane * 4, ring_cn + slot * 128);
            }
        }
    };
    // wait until at most `pending` blocks' copies of this wave are in flight
    auto wait_blocks = [&](int pending) {
        if (COLLECT || pending == 0) {
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        } else if (pending == 1) {
            if constexpr (CPW == 1) asm volatile("s_waitcnt vmcnt(1)" ::: "memory");
            else if constexpr (CPW == 2) asm volatile("s_waitcnt vmcnt(2)" ::: "memory");
            else if constexpr (CPW == 3) asm volatile("s_waitcnt vmcnt(3)" ::: "memory");
            else asm volatile("s_waitcnt vmcnt(5)" ::: "memory");
        } else {
            if constexpr (CPW == 1) asm volatile("s_waitcnt vmcnt(2)" ::: "memory");
            else if constexpr (CPW == 2) asm volatile("s_waitcnt vmcnt(4)" ::: "memory");
            else if constexpr (CPW == 3) asm volatile("s_waitcnt vmcnt(6)" ::: "memory");
            else asm volatile("s_waitcnt vmcnt(10)" ::: "memory");
        }
    };
    static_assert(CPW <= 5, "copy count per wave");

    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t w0 = tile * ROWS_PER_WG + int64_t(wave) * ROWS_PER_WAVE;
        bf16x8 xr[NB][STEPS];
        int64_t rows[NB];
        double T[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int64_t e = w0 + nb * 32 + r32;
            const bool ok = e < nrows;
            int64_t row = e;
            if (COLLECT) {
                row = ok ? int64_t(L.coll_rows[e]) : 0;
                T[nb] = ok ? L.coll_T[e] : -1.0;
            }
            rows[nb] = ok ? row : -1;
            const uint16_t* src = Xb + (ok ? row : 0) * DPB + h * HALF;
#pragma unroll
            for (int s = 0; s < STEPS; ++s) {
                uintx4 v = {0u, 0u, 0u, 0u};
                if (ok) v = *reinterpret_cast<const uintx4*>(src + 8 * s);
                xr[nb][s] = __builtin_bit_cast(bf16x8, v);
            }
        }
        uint32_t m1[NB][NSL], m2[NB][NSL];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int q = 0; q < NSL; ++q) {
                m1[nb][q] = 0xFFFFFFFFu;
                m2[nb][q] = 0xFFFFFFFFu;
            }

        auto block = [&](int slot, int mb, int sl) {
            const float* cnb = reinterpret_cast<const float*>(ring_cn + slot * 128);
            const unsigned char* arow = ring + slot * BLKB + r32 * ROWB;
            // all A fragments and start values of the block first, then the MFMA chain
            bf16x8 a[STEPS];
#pragma unroll
            for (int s = 0; s < STEPS; ++s) {
                const int pc = (h * (HALF / 8) + s) ^ (r32 & SWZ);
                a[s] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uintx4*>(arow + pc * 16));
            }
            floatx16 init;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const floatx4 c4 = *reinterpret_cast<const floatx4*>(cnb + 8 * q + 4 * h);
#pragma unroll
                for (int t = 0; t < 4; ++t) init[4 * q + t] = c4[t];
            }
            floatx16 acc[NB];
#pragma unroll
            for (int s = 0; s < STEPS; ++s) {
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    acc[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[s], xr[nb][s],
                                                                      s == 0 ? init : acc[nb], 0, 0, 0);
            }
            if (COLLECT) {
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                    for (int i = 0; i < 16; ++i) {
                        const uint32_t key = (__float_as_uint(acc[nb][i]) & kmask);
                        if (rows[nb] >= 0 && double(__uint_as_float(key)) <= T[nb]) {
                            const int cj = mb * 32 + 8 * (i >> 2) + 4 * h + (i & 3);
                            if (cj < k) {
                                const int64_t e = w0 + nb * 32 + r32;
                                const int pos = atomicAdd(&L.coll_count[e], 1);
                                if (pos < MAXC2) L.coll_ids[e * MAXC2 + pos] = cj;
                            }
                        }
                    }
            } else {
                const uint32_t tag = uint32_t(mb) << 4;
                const uint32_t lowmask = ~kmask;
                // the mask lives in a VGPR so (bits & mask) | tag is one v_and_or_b32
                // (a VALU op may read only one scalar operand on this target)
                uint32_t kmv;
                asm volatile("v_mov_b32 %0, %1" : "=v"(kmv) : "s"(kmask));
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                    for (int i = 0; i < 16; i += 2) {
                        // bitfield insert: value bits where kmask is set, (block, register) below
                        const uint32_t ta = (tag | uint32_t(i)) & lowmask;
                        const uint32_t tb = (tag | uint32_t(i + 1)) & lowmask;
                        const uint32_t ka = (__float_as_uint(acc[nb][i]) & kmv) | ta;
                        const uint32_t kb = (__float_as_uint(acc[nb][i + 1]) & kmv) | tb;
                        // second smallest of {m1, ka, kb} (m1 <= m2 always)
                        const uint32_t med = max(min(m1[nb][sl], ka), min(max(m1[nb][sl], ka), kb));
                        m2[nb][sl] = min(m2[nb][sl], med);
                        m1[nb][sl] = min(m1[nb][sl], min(ka, kb));
                    }
            }
        };

        // step mb: wait for block mb's copies (blocks mb+1, mb+2 may stay in flight),
        // barrier (all waves' copies visible; everyone is done with block mb-1),
        // refill the slot of block mb-1 with block mb+3, then run block mb
        auto step = [&](int mb, int sl) {
            wait_blocks(min(2, mblocks - 1 - mb));
            __syncthreads();
            if (mb + 3 < mblocks) issue_block(mb + 3, (mb + 3) & (BF16_RING - 1));
            block(mb & (BF16_RING - 1), mb, sl);
        };
        // row loads complete here (a waitcnt the compiler sees, so it does not re-wait
        // for them inside the block loop, where vmcnt also counts the ring copies);
        // the previous tile's last readers are done before the prologue refills the ring
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        __syncthreads();
        issue_block(0, 0);
        if (1 < mblocks) issue_block(1, 1);
        if (2 < mblocks) issue_block(2, 2);
        for (int mb = 0; mb < mblocks; mb += 4) {
            step(mb, 0);
            if (mb + 1 < mblocks) step(mb + 1, 1);
            if (mb + 2 < mblocks) step(mb + 2, 2);
            if (mb + 3 < mblocks) step(mb + 3, 3);
        }
        if (COLLECT) continue;

        // settle rows: lane r32 (h = 0) merges its 4 slots with lane r32 + 32's 4
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            uint32_t k1[2 * NSL], k2[2 * NSL];
#pragma unroll
            for (int q = 0; q < NSL; ++q) {
                k1[q] = m1[nb][q];
                k2[q] = m2[nb][q];
                k1[NSL + q] = __shfl_xor(m1[nb][q], 32);
                k2[NSL + q] = __shfl_xor(m2[nb][q], 32);
            }
            const int64_t row = rows[nb];
            if (h != 0 || row < 0) continue;
            // decode: key low bits = (block << 4) | register; slot half from position
            auto idx_of = [&](uint32_t key, int hh) {
                const uint32_t low = key & ((1u << kc.ib) - 1u);
                const int mbk = int(low >> 4), i = int(low & 15u);
                return mbk * 32 + 8 * (i >> 2) + 4 * hh + (i & 3);
            };
            uint32_t best = 0xFFFFFFFFu;
            int bs = 0;
#pragma unroll
            for (int s = 0; s < 2 * NSL; ++s)
                if (k1[s] < best) {
                    best = k1[s];
                    bs = s;
                }
            const double X = double(xnorm_up[row]), ex = double(enorm_up[row]);
            const double Tr = bf16_threshold(kc, P, X, ex, double(__uint_as_float(best & kmask)));
            bool cert = true;
            int cnt = 0;
            int cand[MAXC];
#pragma unroll
            for (int s = 0; s < 2 * NSL; ++s) {
                if (k2[s] != 0xFFFFFFFFu && double(__uint_as_float(k2[s] & kmask)) <= Tr) cert = false;
                if (k1[s] != 0xFFFFFFFFu && double(__uint_as_float(k1[s] & kmask)) <= Tr) {
                    const int j = idx_of(k1[s], s >= NSL ? 1 : 0);
                    if (j >= k) cert = false;  // a padded centroid inside the band: collect pass decides
                    if (cnt < MAXC) cand[cnt] = j;
                    cnt++;
                }
            }
            if (cert && cnt == 1) {
                labels[row] = idx_of(best, bs >= NSL ? 1 : 0);
            } else if (cert && cnt <= MAXC) {
                const int e = atomicAdd(&L.counters[0], 1);
                L.cand_rows[e] = int32_t(row);
                L.cand_count[e] = cnt;
                for (int c = 0; c < cnt; ++c) L.cand_ids[int64_t(e) * MAXC + c] = cand[c];
            } else {
                const int e = atomicAdd(&L.counters[2], 1);
                L.coll_rows[e] = int32_t(row);
                L.coll_T[e] = Tr;
                L.coll_count[e] = 0;
            }
        }
    }
}

}  // namespace km
