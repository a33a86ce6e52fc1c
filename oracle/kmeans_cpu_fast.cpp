// kmeans_cpu_fast.cpp -- CPU baseline, not the product and not the bit-exact
// oracle: the reference's CPU Lloyd algorithm (DAAL,
// daal/src/algorithms/kmeans/kmeans_lloyd_impl.i:136-246 and
// kmeans_lloyd_batch_impl.i:117-213) with its X.C^T product written as a
// register-blocked, vectorised kernel, so that the CPU figure a benchmark prints
// beside the GPU is what a BLAS-backed build of the reference would reach, not a
// scalar triple loop.
//
// Per row block (one thread each, as addNTToTaskThreadedDense): for every
// centroid panel, goal_j = 0.5 |c_j|^2 - x.c_j (kmeans_lloyd_impl.i:195, xxgemm
// with alpha -1 and beta 1 over the half squared norms), label = the first j
// with the smallest goal (strict <), the row's goal 2 goal_min + |x|^2 added to
// the thread's objective, x added to the thread's float partial sums S1 and S0,
// and the row offered to the thread's list of the k largest goals (empty-cluster
// candidates, kmeans_lloyd_impl.i:413-436).  The threads' S1 are merged in
// double (kmeans_lloyd_impl.i:384-401), empty clusters take the largest
// candidates in ascending cluster order, and the iteration count is fixed
// (the GPU driver's rule with threshold 0).
//
// The product kernel: centroids packed into panels of W (16 for AVX2, 32 for
// AVX-512) feature-major, R rows (6 or 8) by W centroids of accumulators in
// registers, one broadcast of x per row and feature, fused multiply-adds.  The
// instruction set is chosen at run time (AVX-512F, else AVX2 with FMA, else a
// portable loop), so the library built here runs on the GPU host's CPU.
// Rounding differs from the oracle's (fused, reordered sums): the results match
// the reference's algorithm to float rounding, which the tests check.
//
// Build: oracle/Makefile (g++ -O3; the vector kernels carry their own target
// attributes).

#include <immintrin.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

template <typename Fn>
void parallel_ranges(int64_t count, int nthreads, Fn fn) {
    if (nthreads <= 1 || count <= 1) {
        fn(0, int64_t(0), count);
        return;
    }
    if (nthreads > count) nthreads = int(count);
    std::vector<std::thread> pool;
    pool.reserve(nthreads);
    for (int t = 0; t < nthreads; ++t) {
        const int64_t b = count * t / nthreads;
        const int64_t e = count * (t + 1) / nthreads;
        pool.emplace_back([=] { fn(t, b, e); });
    }
    for (auto& th : pool) th.join();
}

// centroids -> panels of W: P[panel][f][w] = c[panel * W + w][f] (0 past k)
void pack_panels(const float* c, int64_t k, int64_t d, int W, std::vector<float>& P) {
    const int64_t np = (k + W - 1) / W;
    P.assign(size_t(np * d * W), 0.f);
    for (int64_t j = 0; j < k; ++j)
        for (int64_t f = 0; f < d; ++f) P[size_t(((j / W) * d + f) * W + (j % W))] = c[j * d + f];
}

// dot[r][w] = x_r . c_{panel, w} for R rows (row stride d), one panel
__attribute__((target("avx512f"))) void panel_dot_avx512(const float* x, int R, int64_t d,
                                                          const float* P, float* dot) {
    __m512 a[8][2];
    for (int r = 0; r < 8; ++r) a[r][0] = a[r][1] = _mm512_setzero_ps();
    if (R == 8) {
        for (int64_t f = 0; f < d; ++f) {
            const __m512 c0 = _mm512_loadu_ps(P + f * 32);
            const __m512 c1 = _mm512_loadu_ps(P + f * 32 + 16);
#pragma GCC unroll 8
            for (int r = 0; r < 8; ++r) {
                const __m512 xb = _mm512_set1_ps(x[r * d + f]);
                a[r][0] = _mm512_fmadd_ps(xb, c0, a[r][0]);
                a[r][1] = _mm512_fmadd_ps(xb, c1, a[r][1]);
            }
        }
    } else {
        for (int64_t f = 0; f < d; ++f) {
            const __m512 c0 = _mm512_loadu_ps(P + f * 32);
            const __m512 c1 = _mm512_loadu_ps(P + f * 32 + 16);
            for (int r = 0; r < R; ++r) {
                const __m512 xb = _mm512_set1_ps(x[r * d + f]);
                a[r][0] = _mm512_fmadd_ps(xb, c0, a[r][0]);
                a[r][1] = _mm512_fmadd_ps(xb, c1, a[r][1]);
            }
        }
    }
    for (int r = 0; r < R; ++r) {
        _mm512_storeu_ps(dot + r * 32, a[r][0]);
        _mm512_storeu_ps(dot + r * 32 + 16, a[r][1]);
    }
}

__attribute__((target("avx2,fma"))) void panel_dot_avx2(const float* x, int R, int64_t d,
                                                        const float* P, float* dot) {
    __m256 a[6][2];
    for (int r = 0; r < 6; ++r) a[r][0] = a[r][1] = _mm256_setzero_ps();
    if (R == 6) {
        for (int64_t f = 0; f < d; ++f) {
            const __m256 c0 = _mm256_loadu_ps(P + f * 16);
            const __m256 c1 = _mm256_loadu_ps(P + f * 16 + 8);
#pragma GCC unroll 6
            for (int r = 0; r < 6; ++r) {
                const __m256 xb = _mm256_broadcast_ss(x + r * d + f);
                a[r][0] = _mm256_fmadd_ps(xb, c0, a[r][0]);
                a[r][1] = _mm256_fmadd_ps(xb, c1, a[r][1]);
            }
        }
    } else {
        for (int64_t f = 0; f < d; ++f) {
            const __m256 c0 = _mm256_loadu_ps(P + f * 16);
            const __m256 c1 = _mm256_loadu_ps(P + f * 16 + 8);
            for (int r = 0; r < R; ++r) {
                const __m256 xb = _mm256_broadcast_ss(x + r * d + f);
                a[r][0] = _mm256_fmadd_ps(xb, c0, a[r][0]);
                a[r][1] = _mm256_fmadd_ps(xb, c1, a[r][1]);
            }
        }
    }
    for (int r = 0; r < R; ++r) {
        _mm256_storeu_ps(dot + r * 16, a[r][0]);
        _mm256_storeu_ps(dot + r * 16 + 8, a[r][1]);
    }
}

void panel_dot_portable(const float* x, int R, int64_t d, const float* P, int W, float* dot) {
    for (int r = 0; r < R; ++r)
        for (int w = 0; w < W; ++w) {
            float s = 0.f;
            for (int64_t f = 0; f < d; ++f) s += x[r * d + f] * P[f * W + w];
            dot[r * W + w] = s;
        }
}

struct kernel_choice {
    int W, R, isa;  // isa: 2 avx512, 1 avx2, 0 portable
};

kernel_choice choose_kernel() {
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx512f")) return {32, 8, 2};
    if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) return {16, 6, 1};
    return {16, 4, 0};
}

struct tls {
    std::vector<int> s0;
    std::vector<float> s1;
    float goal = 0.f;
    size_t cnum = 0;
    std::vector<float> cvalues;
    std::vector<size_t> cindices;
};

// kmeansInsertCandidate: the k largest goals, descending; equal values after
// the ones already kept
void insert_candidate(tls& t, size_t clnum, float value, size_t index) {
    size_t cpos = t.cnum;
    while (cpos > 0 && t.cvalues[cpos - 1] < value) {
        if (cpos < clnum) {
            t.cvalues[cpos] = t.cvalues[cpos - 1];
            t.cindices[cpos] = t.cindices[cpos - 1];
        }
        cpos--;
    }
    if (cpos < clnum) {
        t.cvalues[cpos] = value;
        t.cindices[cpos] = index;
        if (t.cnum < clnum) t.cnum++;
    }
}

}  // namespace

extern "C" {

// Which vector kernel this CPU runs: 2 AVX-512F, 1 AVX2+FMA, 0 portable.
int cpu_fast_isa() { return choose_kernel().isa; }

// `iters` Lloyd iterations from c0 (fixed count), then (final != 0) the final
// assignment and the exact objective sum |x - c_label|^2
// (kmeans_lloyd_postprocessing.h:64-183); without it labels_out holds the last
// iteration's assignment and *obj_out is 0.  Returns 1 when more clusters are
// empty than candidates exist.
int cpu_fast_lloyd_f32(const float* x, int64_t n, int64_t d, const float* c0, int64_t k,
                       int64_t iters, int64_t block_rows, int nthreads, int final, float* c_out,
                       int32_t* labels_out, double* obj_out) {
    const kernel_choice kc = choose_kernel();
    const int W = kc.W, R = kc.R;
    if (block_rows < R) block_rows = R;
    block_rows = block_rows / R * R;
    const int64_t np = (k + W - 1) / W;
    std::vector<float> c(c0, c0 + k * d), P, half(static_cast<size_t>(np * W)), xsq(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        float s = 0.f;
        for (int64_t f = 0; f < d; ++f) s += x[i * d + f] * x[i * d + f];
        xsq[size_t(i)] = s;
    }
    const int64_t nblocks = (n + block_rows - 1) / block_rows;
    const int nthr = int(std::max<int64_t>(1, std::min<int64_t>(nthreads, nblocks)));
    std::vector<int32_t> lab(static_cast<size_t>(n));

    // labels (and goals) of every row against c: the panel loop of one block
    auto assign_block = [&](int64_t row0, int64_t bs, float* goal_out) {
        std::vector<float> dot(static_cast<size_t>(R) * W);
        for (int64_t r0 = 0; r0 < bs; r0 += R) {
            const int rr = int(std::min<int64_t>(R, bs - r0));
            const float* xr = x + (row0 + r0) * d;
            float best[8];
            int32_t bidx[8];
            for (int r = 0; r < rr; ++r) {
                best[r] = std::numeric_limits<float>::infinity();
                bidx[r] = 0;
            }
            for (int64_t p = 0; p < np; ++p) {
                const float* Pp = P.data() + size_t(p * d * W);
                if (kc.isa == 2) panel_dot_avx512(xr, rr, d, Pp, dot.data());
                else if (kc.isa == 1) panel_dot_avx2(xr, rr, d, Pp, dot.data());
                else panel_dot_portable(xr, rr, d, Pp, W, dot.data());
                const int wmax = int(std::min<int64_t>(W, k - p * W));
                for (int r = 0; r < rr; ++r) {
                    const float* dr = dot.data() + r * W;
                    for (int w = 0; w < wmax; ++w) {
                        const float g = half[size_t(p * W + w)] - dr[w];
                        if (g < best[r] || (p == 0 && w == 0)) {
                            best[r] = g;
                            bidx[r] = int32_t(p * W + w);
                        }
                    }
                }
            }
            for (int r = 0; r < rr; ++r) {
                lab[size_t(row0 + r0 + r)] = bidx[r];
                if (goal_out) goal_out[r0 + r] = best[r];
            }
        }
    };
    auto prepare = [&]() {
        pack_panels(c.data(), k, d, W, P);
        std::fill(half.begin(), half.end(), 0.f);
        for (int64_t j = 0; j < k; ++j) {
            float s = 0.f;
            for (int64_t f = 0; f < d; ++f) s += c[size_t(j * d + f)] * c[size_t(j * d + f)];
            half[size_t(j)] = 0.5f * s;
        }
    };

    std::vector<double> ds1(static_cast<size_t>(d));
    for (int64_t it = 0; it < iters; ++it) {
        prepare();
        std::vector<tls> T(static_cast<size_t>(nthr));
        for (auto& t : T) {
            t.s0.assign(size_t(k), 0);
            t.s1.assign(size_t(k * d), 0.f);
            t.cvalues.assign(size_t(k), 0.f);
            t.cindices.assign(size_t(k), 0);
        }
        parallel_ranges(nblocks, nthr, [&](int ti, int64_t b0, int64_t b1) {
            tls& t = T[size_t(ti)];
            std::vector<float> goal(static_cast<size_t>(block_rows));
            for (int64_t blk = b0; blk < b1; ++blk) {
                const int64_t row0 = blk * block_rows;
                const int64_t bs = std::min(block_rows, n - row0);
                assign_block(row0, bs, goal.data());
                float bgoal = 0.f;
                for (int64_t i = 0; i < bs; ++i) {
                    const int64_t row = row0 + i;
                    const int32_t j = lab[size_t(row)];
                    const float g = goal[size_t(i)] * 2.f + xsq[size_t(row)];
                    float* s1 = t.s1.data() + size_t(j) * size_t(d);
                    const float* xr = x + row * d;
                    for (int64_t f = 0; f < d; ++f) s1[f] += xr[f];
                    t.s0[size_t(j)]++;
                    insert_candidate(t, size_t(k), g, size_t(row));
                    bgoal += g;
                }
                t.goal += bgoal;
            }
        });
        // merge: S1 in double, candidates by value (kmeans_lloyd_impl.i:384-480)
        std::vector<float> cv(static_cast<size_t>(k));
        std::vector<size_t> ci(static_cast<size_t>(k));
        size_t cnum = 0;
        {
            std::vector<float> tv(static_cast<size_t>(k));
            std::vector<size_t> tidx(static_cast<size_t>(k));
            for (auto& t : T) {
                size_t cp = 0, lp = 0;
                while (cp + lp < size_t(k) && (cp < cnum || lp < t.cnum)) {
                    if (cp < cnum && (lp == t.cnum || cv[cp] > t.cvalues[lp])) {
                        tv[cp + lp] = cv[cp];
                        tidx[cp + lp] = ci[cp];
                        cp++;
                    } else {
                        tv[cp + lp] = t.cvalues[lp];
                        tidx[cp + lp] = t.cindices[lp];
                        lp++;
                    }
                }
                cnum = cp + lp;
                std::copy(tv.begin(), tv.begin() + long(cnum), cv.begin());
                std::copy(tidx.begin(), tidx.begin() + long(cnum), ci.begin());
            }
        }
        size_t cpos = 0;
        for (int64_t j = 0; j < k; ++j) {
            int s0 = 0;
            for (auto& t : T) s0 += t.s0[size_t(j)];
            if (s0 > 0) {
                std::fill(ds1.begin(), ds1.end(), 0.0);
                for (auto& t : T)
                    for (int64_t f = 0; f < d; ++f) ds1[size_t(f)] += double(t.s1[size_t(j * d + f)]);
                const float coeff = float(1.0 / s0);
                for (int64_t f = 0; f < d; ++f) c[size_t(j * d + f)] = float(ds1[size_t(f)]) * coeff;
            } else {
                if (cpos >= cnum) return 1;
                std::memcpy(&c[size_t(j * d)], x + int64_t(ci[cpos]) * d, sizeof(float) * size_t(d));
                cpos++;
            }
        }
    }
    if (!final) {
        std::copy(c.begin(), c.end(), c_out);
        std::copy(lab.begin(), lab.end(), labels_out);
        *obj_out = 0.0;
        return 0;
    }
    // final assignment and exact objective
    prepare();
    parallel_ranges(nblocks, nthr, [&](int, int64_t b0, int64_t b1) {
        for (int64_t blk = b0; blk < b1; ++blk) {
            const int64_t row0 = blk * block_rows;
            assign_block(row0, std::min(block_rows, n - row0), nullptr);
        }
    });
    std::vector<double> part(static_cast<size_t>(nthr), 0.0);
    parallel_ranges(n, nthr, [&](int ti, int64_t i0, int64_t i1) {
        double s = 0.0;
        for (int64_t i = i0; i < i1; ++i) {
            const float* cr = c.data() + size_t(lab[size_t(i)]) * size_t(d);
            for (int64_t f = 0; f < d; ++f) {
                const double df = double(x[i * d + f]) - double(cr[f]);
                s += df * df;
            }
        }
        part[size_t(ti)] = s;
    });
    double obj = 0.0;
    for (double v : part) obj += v;
    std::copy(c.begin(), c.end(), c_out);
    std::copy(lab.begin(), lab.end(), labels_out);
    *obj_out = obj;
    return 0;
}

}  // extern "C"
