// km_probe.hpp -- measured ceiling of the BF16 matrix cores on this device.
//
// The roofline of the BF16 assignment is quoted against the published dense
// peak.  Under sustained matrix-core load the shader clock runs below the clock
// that peak assumes, so the library also measures what a pure MFMA loop
// reaches: every wave runs four independent v_mfma_f32_32x32x16_bf16 chains
// (no memory traffic inside the loop), and wave 0 of workgroup 0 reads the
// shader-clock and the constant 100 MHz counters around its loop.
#pragma once

#include "km_assign.hpp"
#include "km_assign_bf16.hpp"

namespace km {

constexpr int PROBE_WG_THREADS = 256;

// src: 512 16-byte vectors (bf16 operands); out: one float per thread;
// clk[0], clk[1]: shader-clock and 100 MHz ticks of wave 0 over its loop
__global__ __launch_bounds__(PROBE_WG_THREADS, 2) void mfma_ceiling_kernel(
    const uintx4* __restrict__ src, float* __restrict__ out, long long* __restrict__ clk, int iters) {
    const int t = threadIdx.x;
    const bf16x8 a = __builtin_bit_cast(bf16x8, src[t]);
    const bf16x8 b = __builtin_bit_cast(bf16x8, src[t + PROBE_WG_THREADS]);
    floatx16 acc0 = {0}, acc1 = {0}, acc2 = {0}, acc3 = {0};
    const long long c0 = clock64();
    const long long r0 = wall_clock64();
    for (int it = 0; it < iters; ++it) {
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc1, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc2, 0, 0, 0);
        acc3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc3, 0, 0, 0);
    }
    const long long c1 = clock64();
    const long long r1 = wall_clock64();
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 16; ++i) s += acc0[i] + acc1[i] + acc2[i] + acc3[i];
    out[int64_t(blockIdx.x) * PROBE_WG_THREADS + t] = s;
    if (t == 0 && blockIdx.x == 0) {
        clk[0] = c1 - c0;
        clk[1] = r1 - r0;
    }
}

}  // namespace km
