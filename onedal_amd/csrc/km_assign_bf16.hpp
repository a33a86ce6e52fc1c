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

// Centroid count padded to whole groups of four 32-centroid blocks: the block
// loop of the assignment kernel then runs without tail conditions.  Padding
// centroids are zero rows with a huge start value (never chosen, never inside a
// threshold).
__host__ __device__ inline int bf16_padded_k(int k) { return (k + 127) / 128 * 128; }

// ---------------------------------------------------------------------------
// X -> padded BF16 rows (stride dpb), ||x|| and ||x - bf16(x)|| (rounded up),
// and their maxima.  A row is covered by L = dpb / 4 consecutive lanes, lane f
// converting features 4f .. 4f + 3 (zero past d), so a wave reads 64 / L rows as
// one contiguous run of 16-byte loads and writes their bf16 rows as 8-byte stores;
// the loads of four such steps are issued together.  The per-row sums are FP64
// over shuffles; the 1e-12 factor covers their rounding, so their order is free.
// dpb = pad_dim(d) <= 256 (a power of two, at least 32).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void convert_rows_bf16_kernel(
    const float* __restrict__ X, int64_t n, int d, int dpb, uint16_t* __restrict__ Xb,
    float* __restrict__ xnorm_up, float* __restrict__ enorm_up,
    unsigned long long* __restrict__ max_bits /* [0] max ||x||, [1] max ex */) {
    const int lane = threadIdx.x & 63;
    const int L = dpb >> 2;                 // lanes per row: 8 .. 64
    const int R = 64 / L;                   // rows per wave step
    const int fl = lane & (L - 1);          // this lane's 4-feature group
    const int sub = lane / L;               // this lane's row within the step
    const int64_t wave = (blockIdx.x * int64_t(blockDim.x) + threadIdx.x) >> 6;
    const int64_t nwaves = (int64_t(gridDim.x) * blockDim.x) >> 6;
    const bool vec = (d % 4) == 0 && (reinterpret_cast<uintptr_t>(X) % 16) == 0;
    const int f0 = 4 * fl;
    double mx = 0.0, me = 0.0;
    constexpr int U = 4;  // wave steps whose loads are issued together
    for (int64_t base = wave * R * U; base < n; base += nwaves * R * U) {
        float v[U][4];
#pragma unroll
        for (int st = 0; st < U; ++st) {
            const int64_t row = base + st * R + sub;
#pragma unroll
            for (int u = 0; u < 4; ++u) v[st][u] = 0.f;
            if (row < n) {
                const float* x = X + row * d;
                if (vec && f0 < d) {
                    const float4 t = *reinterpret_cast<const float4*>(x + f0);
                    v[st][0] = t.x;
                    v[st][1] = t.y;
                    v[st][2] = t.z;
                    v[st][3] = t.w;
                } else if (!vec) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) v[st][u] = f0 + u < d ? x[f0 + u] : 0.f;
                }
            }
        }
#pragma unroll
        for (int st = 0; st < U; ++st) {
            const int64_t row = base + st * R + sub;
            const bool ok = row < n;
            double s = 0.0, e = 0.0;
            uint16_t b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                b[u] = f2bf(v[st][u]);
                const double dv = double(v[st][u]);
                const double ev = dv - double(bf2f(b[u]));
                s = fma(dv, dv, s);
                e = fma(ev, ev, e);
            }
            if (ok)
                *reinterpret_cast<uint2*>(Xb + row * dpb + f0) =
                    make_uint2(uint32_t(b[0]) | (uint32_t(b[1]) << 16), uint32_t(b[2]) | (uint32_t(b[3]) << 16));
            for (int off = L >> 1; off > 0; off >>= 1) {
                s += __shfl_xor(s, off);
                e += __shfl_xor(e, off);
            }
            if (ok && fl == 0) {
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
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, off));
        me = fmax(me, __shfl_xor(me, off));
    }
    if (lane == 0) {
        atomicMax(&max_bits[0], (unsigned long long)__double_as_longlong(mx));
        atomicMax(&max_bits[1], (unsigned long long)__double_as_longlong(me));
    }
}

// ---------------------------------------------------------------------------
// centroids -> bf16(-2 c) padded rows, ||c||^2 (F), maxima of ||c||,
// ||bf16(c)||, ||c - bf16(c)||.  One wave per centroid.  Rows k .. kp-1
// (kp = bf16_padded_k(k)) are zero, so the last centroid blocks read defined
// values without clamping.
// max_bits: [0] C, [1] Ch, [2] Ec
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void centroid_prep_bf16_kernel(
    const float* __restrict__ C, int k, int kp, int d, int dpb, uint16_t* __restrict__ Cb,
    float* __restrict__ cn, unsigned long long* __restrict__ max_bits) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    for (int j = k + wave; j < kp; j += nwaves)
        for (int f = lane; f < dpb; f += 64) Cb[int64_t(j) * dpb + f] = 0;
    unsigned long long b0 = 0, b1 = 0, b2 = 0;  // lane 0: this wave's maxima
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
            b0 = max(b0, (unsigned long long)__double_as_longlong(sqrt(s) * up));
            b1 = max(b1, (unsigned long long)__double_as_longlong(sqrt(sh) * up));
            b2 = max(b2, (unsigned long long)__double_as_longlong(sqrt(se) * up + 1e-30));
        }
    }
    // one atomic per bound and workgroup (the maxima of non-negative doubles, or
    // NaN, taken on their bit patterns as before)
    __shared__ unsigned long long wb[4][3];
    const int w = threadIdx.x >> 6;
    if (lane == 0) {
        wb[w][0] = b0;
        wb[w][1] = b1;
        wb[w][2] = b2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int q = threadIdx.x;
        const unsigned long long m = max(max(wb[0][q], wb[1][q]), max(wb[2][q], wb[3][q]));
        if (m) atomicMax(&max_bits[q], m);
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
// a huge start value so they are never chosen.  kp = bf16_padded_k(k) entries.
__global__ void cnb_kernel(const float* __restrict__ cn, int k, int kp,
                           const bf16_params* __restrict__ params_p, float* __restrict__ cnb) {
    const float Bf = float(params_p->B);
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < kp; j += gridDim.x * blockDim.x)
        cnb[j] = j < k ? add_round_up(cn[j], Bf) : __uint_as_float(0x7F000000u);
}

// Centroid blocks live in a ring of LDS slots, filled ahead of use with
// direct-to-LDS loads.  The main kernel uses 4 slots (one barrier per block,
// copies three blocks ahead; 8 slots with a barrier per two blocks measured
// slower).  The collect kernel, whose LDS also holds hit lists, uses 4 slots, or
// 2 for 16 KB blocks (copies one ahead).  A ring of 8 would use one barrier per
// two blocks (bf16 kernel GRP).
__host__ __device__ constexpr int bf16_ring(int dpb, bool collect) {
    return collect ? (dpb <= 128 ? 4 : 2) : 4;
}

// a wave-uniform pointer moved to scalar registers (an SGPR asm operand)
__device__ __forceinline__ const void* uniform_ptr(const void* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
    return reinterpret_cast<const void*>((uint64_t(hi) << 32) | lo);
}

// LDS byte address of a shared-memory pointer, as a wave-uniform value
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
    return __builtin_amdgcn_readfirstlane(uint32_t(reinterpret_cast<uintptr_t>(p)));
}

// 16 bytes per lane from global memory straight into LDS: lane l reads
// gbase + voff (scalar 64-bit base, 32-bit lane offset) and writes
// lds_base + 16 l (wave-uniform, through M0).  Issued as inline asm so the
// compiler does not insert its own conservative waits (a full lgkmcnt wait before
// the block's LDS reads, vmcnt(0) at barriers); ordering is by the explicit vmcnt
// waits and barriers of the caller.  M0 is an input operand: the compiler writes
// it before the asm and knows the asm leaves it unchanged.
// The compiler's hazard recognizer does not look inside inline asm, so the asm
// avoids the hazards itself: the base SGPRs typically come from v_readfirstlane
// just before, and a VMEM read of an SGPR written by a VALU op needs 5 wait
// states (otherwise the load may use the register's previous contents as the
// address).  The load therefore reads a copy made by a scalar move (no wait
// states after an SALU write), and that move is also the one wait state that an
// M0 write -> LDS DMA needs.
__device__ __forceinline__ void lds_dma16(const void* gbase, uint32_t voff, uint32_t lds_base) {
    uint64_t base;
    asm volatile(
        "s_mov_b64 %0, %3\n\t"
        "global_load_lds_dwordx4 %1, %0"
        : "=&s"(base)
        : "v"(voff), "{m0}"(lds_base), "s"(gbase)
        : "memory");
}

// wait until at most N vector-memory operations of this wave are outstanding
template <int N>
__device__ __forceinline__ void wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// 4 bytes per lane, as lds_dma16 (same hazard handling): lane l writes lds_base + 4 l
__device__ __forceinline__ void lds_dma4(const void* gbase, uint32_t voff, uint32_t lds_base) {
    uint64_t base;
    asm volatile(
        "s_mov_b64 %0, %3\n\t"
        "global_load_lds_dword %1, %0"
        : "=&s"(base)
        : "v"(voff), "{m0}"(lds_base), "s"(gbase)
        : "memory");
}

// Re-score lists of the main kernel, one segment of `cap` entries per wave
// (segment = 4 * workgroup + wave); compact_lists_kernel joins them.
struct seg_lists {
    int32_t* cand_rows;
    int32_t* cand_count;
    int32_t* cand_ids;   // MAXC per entry
    int32_t* coll_rows;
    double* coll_T;
    int32_t* counts;     // [2 * segment]: candidate entries, [2 * segment + 1]: collect entries
    int64_t cap;
};

// number of set bits of mask in lanes below this one
__device__ __forceinline__ int lanes_below(uint64_t mask) {
    return int(__builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u)));
}

// LDS bytes of assign_bf16_kernel: ring of 32 x dpb bf16 blocks, and per ring
// slot 256 bytes holding the block's 32 cnB values twice; the collect kernel
// adds per-row hit lists (a count and MAXC2 ids per tile row).  Every layout stays
// below 68 KB (67584 bytes at most).
__host__ inline size_t bf16_lds_bytes(int dpb) {
    return size_t(bf16_ring(dpb, false)) * (size_t(32) * dpb * 2 + 256);
}
__host__ inline size_t bf16_collect_lds_bytes(int dpb, int rows_per_wg) {
    return size_t(bf16_ring(dpb, true)) * (size_t(32) * dpb * 2 + 256) +
           size_t(rows_per_wg) * (MAXC2 + 1) * 4;
}

// ---------------------------------------------------------------------------
// Main / collect kernel.  Workgroup = 4 waves; each wave owns NB blocks of 32
// rows (B operand in VGPRs, 8 bf16 per step); centroid blocks of 32 rows (A
// operand) stream through an LDS ring by global_load_lds_dwordx4 (bf16_ring).  A block is stored unpadded, 16-byte chunk c of row r at chunk
// c ^ (r & SWZ), so the 32 lanes reading one chunk column hit distinct banks.
// Lane l = 32*h + r: half h covers features [h*DPB/2, (h+1)*DPB/2), 8 per step.
// Slots: lane keeps 4 (centroid block index mod 4); a row has 8.
// The key work of block mb runs while the MFMAs of block mb+1 execute (two
// accumulator sets), interleaved by the compiler.
// Cb and cnb_g hold bf16_padded_k(k) rows, a multiple of four blocks.
// COLLECT: rows come from L.coll_rows; every j with v_j <= T is recorded in
// per-row LDS lists (the wave that owns a row is the only writer, so no
// barrier guards them) and flushed to L.coll_count / L.coll_ids once per tile.
// ---------------------------------------------------------------------------
template <int DPB, int NB, bool COLLECT>
__global__ __launch_bounds__(256, 2) void assign_bf16_kernel(
    const uint16_t* __restrict__ Xb, int64_t n, const uint16_t* __restrict__ Cb, int k,
    const float* __restrict__ cnb_g, const float* __restrict__ xnorm_up,
    const float* __restrict__ enorm_up, const bf16_params* __restrict__ params_p,
    bf16_consts kc, int32_t* __restrict__ labels, refine_lists L, seg_lists S) {
    constexpr int HALF = DPB / 2;
    constexpr int STEPS = HALF / 8;
    constexpr int ROWS_PER_WAVE = 32 * NB;
    constexpr int ROWS_PER_WG = 4 * ROWS_PER_WAVE;
    constexpr int CPR = DPB / 8;                  // 16-byte chunks per centroid row
    constexpr int SWZ = (CPR < 16 ? CPR : 16) - 1;  // chunk swizzle mask
    constexpr int ROWB = DPB * 2;                 // bytes per centroid row
    constexpr int BLKB = 32 * ROWB;               // bytes per block
    constexpr int NDATA = BLKB / 1024;            // 1 KB copies per block
    constexpr int DPW = NDATA >= 4 ? NDATA / 4 : 1;  // data copies per wave per block
    constexpr int CPW = DPW + 1;                  // + the wave's quarter of the block's cnB
    constexpr int NSL = 4;
    constexpr int RING = bf16_ring(DPB, COLLECT);  // LDS slots
    constexpr int GRP = RING >= 8 ? 2 : 1;         // blocks per barrier
    constexpr int AHEAD = RING - GRP;              // blocks copied ahead of use
    static_assert(NDATA >= 2 && (NDATA < 4 || NDATA % 4 == 0), "block copy layout");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned char* ring = smem_raw;                                 // [RING][BLKB]
    unsigned char* ring_cn = smem_raw + size_t(RING) * BLKB;   // [RING][64] floats
    // collect kernel: hit count and ids of each tile row (local row = wave, block, lane)
    int* hit_cnt = reinterpret_cast<int*>(ring_cn + RING * 256);  // [ROWS_PER_WG]
    int* hit_ids = hit_cnt + ROWS_PER_WG;                                // [ROWS_PER_WG][MAXC2]

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int r32 = lane & 31;
    const int h = lane >> 5;
    const int mblocks = bf16_padded_k(k) / 32;  // a multiple of 4
    const bf16_params P = *params_p;
    const uint32_t kmask = 0x7FFFFFFFu & ~((1u << kc.ib) - 1u);
    const int64_t nrows = COLLECT ? int64_t(L.counters[2]) : n;
    const int64_t ntiles = (nrows + ROWS_PER_WG - 1) / ROWS_PER_WG;
    // the mask lives in a VGPR so (bits & mask) | tag is one v_and_or_b32
    // (a VALU op may read only one scalar operand on this target)
    uint32_t kmv;
    asm volatile("v_mov_b32 %0, %1" : "=v"(kmv) : "s"(kmask));

    // This wave's direct-to-LDS copies of a block.  Data copy c fills LDS bytes
    // [q KB, (q+1) KB) of the block, q = wave + 4 c (blocks of 2 KB: both pairs
    // of waves write the same two copies, the same bytes); lane l moves 16-byte
    // chunk p = 64 q + l, i.e. physical row p / CPR, physical chunk p % CPR, which
    // holds logical chunk (p % CPR) ^ (row & SWZ).  Then every wave moves the
    // block's 32 cnB values into the slot's 256 bytes (lane l: value l & 31; all
    // four waves write the same bytes, all lanes active, so the block loop has no
    // divergent branch).  Every wave issues CPW copies per block, so the vmcnt
    // waits are uniform.
    uint32_t voff[DPW], moff[DPW];
#pragma unroll
    for (int c = 0; c < DPW; ++c) {
        const int q = NDATA >= 4 ? wave + 4 * c : wave % NDATA;
        const int p = q * 64 + lane;
        const int prow = p / CPR, pc = p % CPR;
        const int lc = pc ^ (prow & SWZ);
        voff[c] = uint32_t((prow * DPB + lc * 8) * 2);
        moff[c] = lds_addr(ring + q * 1024);
    }
    const uint32_t voff_cn = uint32_t((lane & 31) * 4);
    const uint32_t moff_cn = lds_addr(ring_cn);
    auto issue_block = [&](int mb, int slot) {
        const void* gb = uniform_ptr(reinterpret_cast<const unsigned char*>(Cb) + size_t(mb) * BLKB);
#pragma unroll
        for (int c = 0; c < DPW; ++c) lds_dma16(gb, voff[c], moff[c] + slot * BLKB);
        const void* gc = uniform_ptr(cnb_g + mb * 32);
        lds_dma4(gc, voff_cn, moff_cn + slot * 256);
    };
    // Copies run AHEAD blocks ahead, also past the last block (those re-copy the
    // last block into a slot nobody reads again in this tile), GRP at a time, so at
    // every barrier the copies of exactly AHEAD - GRP later blocks may stay in
    // flight (the block loop issues no other vector-memory operation: collect hits
    // go to LDS).
    auto wait_block = [&]() { wait_vm<(AHEAD - GRP) * CPW>(); };

    // this wave's list segment (main kernel): at most cap rows pass through a wave
    const int seg = int(blockIdx.x) * 4 + wave;
    const int64_t seg_base = int64_t(seg) * S.cap;
    int n_cand = 0, n_coll = 0;
    // rows of a tile: B operand fragments, ids, and the per-row threshold inputs
    bf16x8 xr[NB][STEPS];
    int64_t rows[NB];
    uint32_t TKc[NB];  // collect: value(key) <= T  <=>  key <= TKc (0: inactive row)
    float xn[NB], xe[NB];  // ||x||, ||x - bf16(x)|| of the row (for the threshold)
    auto load_tile = [&](int64_t tile) {
        const int64_t w0 = tile * ROWS_PER_WG + int64_t(wave) * ROWS_PER_WAVE;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int64_t e = w0 + nb * 32 + r32;
            const bool ok = e < nrows;
            int64_t row = e;
            if (COLLECT) {
                row = ok ? int64_t(L.coll_rows[e]) : 0;
                // keys are positive-float bits with the low IB bits clear, so
                // value(key) <= T  <=>  key <= bits(largest float <= T); a NaN or
                // negative T collects nothing (the row then goes to the full list)
                uint32_t tkc = 0u;
                if (ok) {
                    const double T = L.coll_T[e];
                    if (T >= 0.0) {
                        float tf = float(T);
                        if (double(tf) > T) tf = __uint_as_float(__float_as_uint(tf) - 1u);
                        tkc = __float_as_uint(tf);
                    }
                }
                TKc[nb] = tkc;
            } else {
                xn[nb] = ok ? xnorm_up[row] : 0.f;
                xe[nb] = ok ? enorm_up[row] : 0.f;
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
    };
    if (int64_t(blockIdx.x) < ntiles) load_tile(blockIdx.x);
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t w0 = tile * ROWS_PER_WG + int64_t(wave) * ROWS_PER_WAVE;
        uint32_t m1[NB][NSL], m2[NB][NSL];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int q = 0; q < NSL; ++q) {
                m1[nb][q] = 0xFFFFFFFFu;
                m2[nb][q] = 0xFFFFFFFFu;
            }

        // distances of the block in ring slot `slot`: all A fragments and start
        // values first, then the MFMA chain
        auto mfma_block = [&](int slot, floatx16(&acc)[NB]) {
            const float* cnb = reinterpret_cast<const float*>(ring_cn + slot * 256);
            const unsigned char* arow = ring + slot * BLKB + r32 * ROWB;
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
#pragma unroll
            for (int s = 0; s < STEPS; ++s) {
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    acc[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[s], xr[nb][s],
                                                                      s == 0 ? init : acc[nb], 0, 0, 0);
            }
        };
        // keys of block mb into slot sl (running top-2), or the collect test
        auto keys_block = [&](const floatx16(&acc)[NB], int mb, int sl) {
            if (COLLECT) {
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                    for (int i = 0; i < 16; ++i) {
                        const uint32_t key = (__float_as_uint(acc[nb][i]) & kmask);
                        if (key <= TKc[nb]) {
                            const int cj = mb * 32 + 8 * (i >> 2) + 4 * h + (i & 3);
                            if (cj < k) {
                                const int lr = wave * ROWS_PER_WAVE + nb * 32 + r32;
                                const int pos = atomicAdd(&hit_cnt[lr], 1);
                                if (pos < MAXC2) hit_ids[lr * MAXC2 + pos] = cj;
                            }
                        }
                    }
            } else {
                // tag = (block, register) < 2^IB, since IB = 4 + ceil(log2(blocks))
                const uint32_t tag = uint32_t(mb) << 4;
                // register pairs outermost: each (block, register) tag is live briefly
#pragma unroll
                for (int i = 0; i < 16; i += 2) {
                    // tags formed on the scalar unit, so a key is one v_and_or_b32;
                    // s_or_b32 writes SCC, which the compiler must not keep a
                    // branch condition in across this asm (hence the clobber)
                    uint32_t ta, tb;
                    asm("s_or_b32 %0, %1, %2" : "=s"(ta) : "s"(tag), "i"(i) : "scc");
                    asm("s_or_b32 %0, %1, %2" : "=s"(tb) : "s"(tag), "i"(i + 1) : "scc");
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb) {
                        // value bits where kmask is set, (block, register) below
                        const uint32_t ka = (__float_as_uint(acc[nb][i]) & kmv) | ta;
                        const uint32_t kb = (__float_as_uint(acc[nb][i + 1]) & kmv) | tb;
                        // second smallest of {m1, ka, kb} (m1 <= m2 always)
                        const uint32_t med = max(min(m1[nb][sl], ka), min(max(m1[nb][sl], ka), kb));
                        m2[nb][sl] = min(m2[nb][sl], med);
                        m1[nb][sl] = min(m1[nb][sl], min(ka, kb));
                    }
                }
            }
        };
        // step mb (u = mb mod 4): at the first block of a group, wait for the
        // group's copies (the next AHEAD - GRP blocks may stay in flight), barrier
        // (all waves' copies visible; everyone is done with the previous group),
        // refill the previous group's slots with blocks mb+AHEAD .. ; then the
        // MFMAs of block mb (ring slot mb mod RING) beside the keys of block mb-1
        // (top-2 slot (mb-1) mod 4)
        floatx16 acc0[NB], acc1[NB];
        auto step = [&](int mb, int u, floatx16(&cur)[NB], floatx16(&prev)[NB]) {
            if (u % GRP == 0) {
                // nothing is scheduled across the wait, the barrier and the copies
                __builtin_amdgcn_sched_barrier(0);
                wait_block();
                __syncthreads();
#pragma unroll
                for (int g = 0; g < GRP; ++g)
                    issue_block(min(mb + AHEAD + g, mblocks - 1), (mb + AHEAD + g) & (RING - 1));
                __builtin_amdgcn_sched_barrier(0);
            }
            // the compiler interleaves the key work of block mb - 1 with the
            // MFMAs of block mb (explicit scheduling groups measured slower)
            mfma_block(mb & (RING - 1), cur);
            keys_block(prev, mb - 1, (u + 3) & 3);
        };
        // row loads complete here (a waitcnt the compiler sees, so it does not re-wait
        // for them inside the block loop, where vmcnt also counts the ring copies);
        // the previous tile's last readers are done before the prologue refills the ring
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        if (COLLECT) {
            // this wave's rows start the tile with empty hit lists (the previous
            // tile's flush, by this wave, read them before in LDS order)
            for (int r = lane; r < ROWS_PER_WAVE; r += 64) hit_cnt[wave * ROWS_PER_WAVE + r] = 0;
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < AHEAD; ++b) issue_block(min(b, mblocks - 1), b);
        // the first step's "previous block" (-1): all-ones accumulators give keys
        // 0xFFFFFFFx, above every real key (each slot receives real keys later,
        // which displace them) and NaN-valued, so never below a threshold
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int i = 0; i < 16; ++i) acc1[nb][i] = __uint_as_float(0xFFFFFFFFu);
        for (int mb = 0; mb < mblocks; mb += 4) {
            step(mb, 0, acc0, acc1);
            step(mb + 1, 1, acc1, acc0);
            step(mb + 2, 2, acc0, acc1);
            step(mb + 3, 3, acc1, acc0);
        }
        keys_block(acc1, mblocks - 1, 3);
        // the next tile's rows load while this tile settles (the fragments are dead)
        int64_t rows_s[NB];
        float xn_s[NB], xe_s[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            rows_s[nb] = rows[nb];
            xn_s[nb] = xn[nb];
            xe_s[nb] = xe[nb];
        }
        if (tile + gridDim.x < ntiles) load_tile(tile + gridDim.x);
        if (COLLECT) {
            // flush this wave's hit lists: count (may exceed MAXC2: overflow) and
            // the first MAXC2 ids of each row, row after row
            for (int idx = lane; idx < ROWS_PER_WAVE * (MAXC2 + 1); idx += 64) {
                const int r = idx / (MAXC2 + 1), q = idx - r * (MAXC2 + 1);
                const int64_t e = w0 + r;
                if (e < nrows) {
                    const int lr = wave * ROWS_PER_WAVE + r;
                    const int c = hit_cnt[lr];
                    if (q == MAXC2)
                        L.coll_count[e] = c;
                    else if (q < c)
                        L.coll_ids[e * MAXC2 + q] = hit_ids[lr * MAXC2 + q];
                }
            }
            continue;
        }

        // settle: with two row blocks, lane (r32, h) settles row r32 of block h (both
        // halves busy), with its own 4 slots of that block and the partner lane's 4
        // (one shuffle); with one block the lanes h = 0 settle.  Rows that need a
        // re-score go to this wave's own segment of the lists, at positions counted
        // by ballots (no global atomics).
        {
            const int nbs = NB == 2 ? h : 0;
            uint32_t k1[2 * NSL], k2[2 * NSL];  // own half's slots, then the partner's
#pragma unroll
            for (int q = 0; q < NSL; ++q) {
                k1[q] = nbs ? m1[NB - 1][q] : m1[0][q];
                k2[q] = nbs ? m2[NB - 1][q] : m2[0][q];
                const uint32_t s1 = (NB == 2 && !h) ? m1[NB - 1][q] : m1[0][q];
                const uint32_t s2 = (NB == 2 && !h) ? m2[NB - 1][q] : m2[0][q];
                k1[NSL + q] = __shfl_xor(s1, 32);
                k2[NSL + q] = __shfl_xor(s2, 32);
            }
            const int64_t row = nbs ? rows_s[NB - 1] : rows_s[0];
            const float xnr = nbs ? xn_s[NB - 1] : xn_s[0];
            const float xer = nbs ? xe_s[NB - 1] : xe_s[0];
            const bool active = (NB == 2 || h == 0) && row >= 0;
            // decode: key low bits = (block << 4) | register; hh = half that holds it
            auto idx_of = [&](uint32_t key, int hh) {
                const uint32_t low = key & ((1u << kc.ib) - 1u);
                const int mbk = int(low >> 4), i = int(low & 15u);
                return mbk * 32 + 8 * (i >> 2) + 4 * hh + (i & 3);
            };
            uint32_t bo = k1[0], bp = k1[NSL];
#pragma unroll
            for (int q = 1; q < NSL; ++q) {
                bo = min(bo, k1[q]);
                bp = min(bp, k1[NSL + q]);
            }
            const uint32_t best = min(bo, bp);
            const int hb = bo <= bp ? h : 1 - h;
            const double Tr = bf16_threshold(kc, P, double(xnr), double(xer),
                                             double(__uint_as_float(best & kmask)));
            // value(key) <= Tr  <=>  key <= TK: the threshold rounded down to a float,
            // with the index bits set (positive floats order like their bits; the
            // all-ones sentinels stay above)
            float trf = float(Tr);
            if (double(trf) > Tr) trf = __uint_as_float(__float_as_uint(trf) - 1u);
            const uint32_t TK = (__float_as_uint(trf) & kmask) | ((1u << kc.ib) - 1u);
            // padding centroids have keys >= 0x7F000000: a threshold that high cannot
            // certify (the collect pass decides)
            bool cert = TK < 0x7F000000u;
            int cnt = 0;
#pragma unroll
            for (int q = 0; q < 2 * NSL; ++q) {
                cert = cert && k2[q] > TK;
                cnt += k1[q] <= TK ? 1 : 0;
                // an id past k (a padding centroid) is never certified; such a row
                // goes to the collect pass, which records ids below k only
                cert = cert && (k1[q] > TK || idx_of(k1[q], q < NSL ? h : 1 - h) < k);
            }
            int outcome = 0;  // 1: candidate entry, 2: collect entry
            if (active) {
                if (cert && cnt == 1)
                    labels[row] = idx_of(best, hb);
                else
                    outcome = cert ? 1 : 2;
            }
            const uint64_t mc = __ballot(outcome == 1);
            const uint64_t ml = __ballot(outcome == 2);
            if (outcome == 1) {
                const int64_t e = seg_base + n_cand + lanes_below(mc);
                S.cand_rows[e] = int32_t(row);
                S.cand_count[e] = cnt;
                int c = 0;
#pragma unroll
                for (int q = 0; q < 2 * NSL; ++q)
                    if (k1[q] <= TK) S.cand_ids[e * MAXC + c++] = idx_of(k1[q], q < NSL ? h : 1 - h);
            } else if (outcome == 2) {
                const int64_t e = seg_base + n_coll + lanes_below(ml);
                S.coll_rows[e] = int32_t(row);
                S.coll_T[e] = Tr;
            }
            n_cand += int(__popcll(mc));
            n_coll += int(__popcll(ml));
        }
    }
    if (!COLLECT && lane == 0) {
        S.counts[2 * seg] = n_cand;
        S.counts[2 * seg + 1] = n_coll;
    }
}

// Concatenates the per-wave segments of the main kernel into the dense candidate
// and collect lists (one wave per segment; the order of entries is not significant:
// each entry re-scores its own row).
__global__ __launch_bounds__(256) void compact_lists_kernel(seg_lists S, int nseg, refine_lists L) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int seg = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    // one atomic per list and workgroup: thread 0 reserves the four segments' runs
    // (segments past nseg count zero)
    __shared__ int wbase[4][2];
    if (threadIdx.x == 0) {
        const int s0 = blockIdx.x * 4;
        int c[4], l[4], tc = 0, tl = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            c[q] = s0 + q < nseg ? S.counts[2 * (s0 + q)] : 0;
            l[q] = s0 + q < nseg ? S.counts[2 * (s0 + q) + 1] : 0;
            tc += c[q];
            tl += l[q];
        }
        int bc = tc ? atomicAdd(&L.counters[0], tc) : 0;
        int bl = tl ? atomicAdd(&L.counters[2], tl) : 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            wbase[q][0] = bc;
            wbase[q][1] = bl;
            bc += c[q];
            bl += l[q];
        }
    }
    __syncthreads();
    if (seg >= nseg) return;
    const int nc = S.counts[2 * seg], nl = S.counts[2 * seg + 1];
    const int bc = wbase[wave][0], bl = wbase[wave][1];
    const int64_t base = int64_t(seg) * S.cap;
    // U entries per lane in flight; an entry's MAXC id slots move as two 16-byte
    // vectors (entries are 32-byte aligned; slots past its count are never read)
    constexpr int U = 4;
    static_assert(MAXC == 8, "an entry's ids are two 16-byte vectors");
    for (int i0 = lane; i0 < nc; i0 += 64 * U) {
        int32_t row[U], cnt[U];
        uintx4 id0[U], id1[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + 64 * u;
            const int64_t src = base + (i < nc ? i : 0);
            row[u] = S.cand_rows[src];
            cnt[u] = S.cand_count[src];
            id0[u] = *reinterpret_cast<const uintx4*>(S.cand_ids + src * MAXC);
            id1[u] = *reinterpret_cast<const uintx4*>(S.cand_ids + src * MAXC + 4);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + 64 * u;
            if (i < nc) {
                const int64_t dst = int64_t(bc) + i;
                L.cand_rows[dst] = row[u];
                L.cand_count[dst] = cnt[u];
                *reinterpret_cast<uintx4*>(L.cand_ids + dst * MAXC) = id0[u];
                *reinterpret_cast<uintx4*>(L.cand_ids + dst * MAXC + 4) = id1[u];
            }
        }
    }
    for (int i0 = lane; i0 < nl; i0 += 64 * U) {
        int32_t row[U];
        double T[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + 64 * u;
            const int64_t src = base + (i < nl ? i : 0);
            row[u] = S.coll_rows[src];
            T[u] = S.coll_T[src];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + 64 * u;
            if (i < nl) {
                L.coll_rows[int64_t(bl) + i] = row[u];
                L.coll_T[int64_t(bl) + i] = T[u];
            }
        }
    }
}

}  // namespace km
