// kmeans_oracle.cpp -- CPU oracle for the dense k-means Lloyd path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
// product library (libkmeans_hip.so) never links, loads or calls it.
//
// Two restatements live here, both in plain C++ (no BLAS, no TBB):
//
// 1. daal_lloyd<F>: the reference's CPU Lloyd (DAAL KMeansBatchKernel<lloydDense>)
//    restated step by step:
//      - loop / update / empty-cluster fill / stop rule
//          cpp/daal/src/algorithms/kmeans/kmeans_lloyd_batch_impl.i:117-213
//      - 0.5*|c|^2, GEMM-form goal clSq - x.c, strict '<' argmin, goal = 2*min + |x|^2,
//        per-thread S0/S1 in F, per-thread candidate lists
//          kmeans_lloyd_impl.i:65-78, 136-246, 413-436
//      - S1 merged over threads in double, stored back in F
//          kmeans_lloyd_impl.i:363-411, kmeans_lloyd_batch_impl.i:98-102
//      - candidate merge across threads  kmeans_lloyd_impl.i:438-480
//      - final assignment (alpha=1 GEMM, strict '>' update) and exact objective
//          kmeans_lloyd_postprocessing.h:64-183
//    The X.C^T product is a plain dot product in F: the reference delegates it to
//    MKL/OpenBLAS xxgemm (service_blas.h:137-141), whose summation order is not
//    pinned, so this restatement matches the reference to rounding, not bitwise.
//    The block size is a parameter (the reference derives it from cache sizes in
//    kmeans_lloyd_helper.h:119-146); threads take contiguous block ranges.
//    Stop rules: 0 = DAAL centroid shift (l2Norm <= thr, batch_impl.i:191-202),
//    1 = oneDAL GPU objective rule (thr > 0 && obj + thr > prev,
//    cpp/oneapi/dal/algo/kmeans/backend/gpu/train_kernel_lloyd_dense_dpc.cpp:162-166),
//    2 = run exactly max_iter iterations.
//
// 2. canonical_lloyd<F>: the arithmetic contract of the MI355X engine, stated
//    on the CPU so the engine can be checked bit for bit:
//      - labels: argmin_j of the canonical distance
//            D_j = sum_f (x_f - c_jf)^2   in F, features in order, mul and add
//            rounded separately (no FMA), strict '<' so ties go to the lowest j.
//        This is the brute force the reference's assignment unit test checks
//        labels against (backend/gpu/test/cluster_updater_dpc.cpp:288-333).
//      - per-point distance for the objective and for empty-cluster candidates:
//            d64_i = sum_f ((double)x_f - (double)c_f)^2  in double, same order.
//      - centroid sums in double in a fixed order: the members of cluster j in
//        ascending row order are cut into chunks of CH members; inside a chunk the
//        members are dealt round-robin to G interleaved partial sums (member m goes
//        to partial m % G), each summed sequentially, then the G partials are added
//        in order g = 0..G-1; the chunk totals of a cluster are added in chunk order.
//        C_j = (F)(S1_j / count_j).
//      - empty clusters (count 0), ascending j, take the points with the largest
//        d64 (ties: lowest row index) in that order; the iteration objective drops by
//        their d64.  Same policy as the reference
//        (empty_cluster_handling_dpc.cpp:26-75, 402-455; kmeans_lloyd_batch_impl.i:169-186).
//      - objective: rows cut into chunks of OBJ_CH; in a chunk, partial t sums rows
//        t, t+256, ... in order, the 256 partials are combined by the tree
//        p[t] += p[t+s] (s = 128..1); chunk sums are added in chunk order.
//      - loop and stop rule of the oneDAL GPU driver
//        (train_kernel_lloyd_dense_dpc.cpp:147-196): per-iteration objective is the
//        pre-update objective plus the empty-cluster correction; a final assignment
//        with the final centroids gives labels and the returned objective.
//
// 2b. canonical_lloyd_csr: the same contract for CSR rows (method lloyd_csr), see
//     the section of that name below.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off).  The -ffp-contract=off flag
// matters: every mul/add in this file must round on its own.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <thread>
#include <map>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// threading helper: contiguous ranges of [0, count) over nthreads threads
// ---------------------------------------------------------------------------
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

// ---------------------------------------------------------------------------
// 1. DAAL Lloyd restatement
// ---------------------------------------------------------------------------
template <typename F>
struct daal_tls {
    std::vector<int> s0;
    std::vector<F> s1;
    F goal = F(0);
    size_t cnum = 0;
    std::vector<F> cvalues;
    std::vector<size_t> cindices;
    std::vector<F> xclusters;
};

// kmeans_lloyd_impl.i:413-436 (kmeansInsertCandidate): keeps the clNum largest
// goal values in descending order; an equal value is placed after existing ones.
template <typename F>
void daal_insert_candidate(daal_tls<F>& tt, size_t clnum, F value, size_t index) {
    size_t cpos = tt.cnum;
    while (cpos > 0 && tt.cvalues[cpos - 1] < value) {
        if (cpos < clnum) {
            tt.cvalues[cpos] = tt.cvalues[cpos - 1];
            tt.cindices[cpos] = tt.cindices[cpos - 1];
        }
        cpos--;
    }
    if (cpos < clnum) {
        tt.cvalues[cpos] = value;
        tt.cindices[cpos] = index;
        if (tt.cnum < clnum) tt.cnum++;
    }
}

template <typename F>
void daal_cluster_halfsq(const F* c, int64_t k, int64_t p, std::vector<F>& clsq) {
    clsq.assign(k, F(0));
    for (int64_t j = 0; j < k; ++j) {
        F sum(0);
        for (int64_t f = 0; f < p; ++f) sum += c[j * p + f] * c[j * p + f];
        clsq[j] = sum * F(0.5);
    }
}

// returns 0 on success, 1 if the candidates ran out (ErrorKMeansNumberOfClustersIsTooLarge)
template <typename F>
int daal_lloyd(const F* x, int64_t n, int64_t p, const F* c0, int64_t k, int64_t max_iter,
               double thr, int stop_rule, int64_t block_size, int nthreads, F* c_out,
               int32_t* labels_out, double* obj_out, int64_t* iters_out) {
    std::vector<F> in_clusters(c0, c0 + k * p);
    std::vector<F> clusters(k * p);
    std::vector<int> cluster_s0(k);
    std::vector<F> cluster_s1(k * p);
    std::vector<double> ds1(p);
    std::vector<F> cvalues(k);
    std::vector<size_t> cindices(k);
    const int64_t nblocks = (n + block_size - 1) / block_size;
    if (nthreads < 1) nthreads = 1;
    const int nthr = int(std::min<int64_t>(nthreads, nblocks));
    F old_target(0);
    F prev_obj_gpu_rule = std::numeric_limits<F>::max();
    int64_t kiter = 0;
    std::vector<F> clsq;
    int status = 0;

    for (kiter = 0; kiter < max_iter; kiter++) {
        daal_cluster_halfsq(in_clusters.data(), k, p, clsq);
        std::vector<daal_tls<F>> tls(nthr);
        for (auto& tt : tls) {
            tt.s0.assign(k, 0);
            tt.s1.assign(k * p, F(0));
            tt.goal = F(0);
            tt.cnum = 0;
            tt.cvalues.assign(k, F(0));
            tt.cindices.assign(k, 0);
            tt.xclusters.assign(size_t(block_size) * size_t(k), F(0));
        }
        // addNTToTaskThreadedDense (kmeans_lloyd_impl.i:136-246)
        parallel_ranges(nblocks, nthr, [&](int t, int64_t b0, int64_t b1) {
            daal_tls<F>& tt = tls[t];
            for (int64_t blk = b0; blk < b1; ++blk) {
                const int64_t row0 = blk * block_size;
                const int64_t bs = std::min(block_size, n - row0);
                const F* data = x + row0 * p;
                F* xc = tt.xclusters.data();
                for (int64_t j = 0; j < k; ++j)
                    for (int64_t i = 0; i < bs; ++i) xc[i + j * bs] = clsq[j];
                // xxgemm('t','n', alpha=-1, beta=1): xc = -1 * X.C^T + xc
                for (int64_t j = 0; j < k; ++j) {
                    const F* cj = in_clusters.data() + j * p;
                    for (int64_t i = 0; i < bs; ++i) {
                        F dot(0);
                        for (int64_t f = 0; f < p; ++f) dot += data[i * p + f] * cj[f];
                        xc[i + j * bs] = F(-1) * dot + xc[i + j * bs];
                    }
                }
                F goal(0);
                for (int64_t i = 0; i < bs; ++i) {
                    F min_goal = xc[i];
                    int64_t min_idx = 0;
                    for (int64_t j = 0; j < k; ++j) {
                        const F v = xc[i + j * bs];
                        if (v < min_goal) {
                            min_goal = v;
                            min_idx = j;
                        }
                    }
                    min_goal *= F(2.0);
                    for (int64_t f = 0; f < p; ++f) {
                        tt.s1[min_idx * p + f] += data[i * p + f];
                        min_goal += data[i * p + f] * data[i * p + f];
                    }
                    daal_insert_candidate(tt, size_t(k), min_goal, size_t(row0 + i));
                    tt.s0[min_idx]++;
                    goal += min_goal;
                }
                tt.goal += goal;
            }
        });
        // kmeansComputeCentroids with double auxData (kmeans_lloyd_impl.i:384-401)
        for (int64_t j = 0; j < k; ++j) {
            std::fill(ds1.begin(), ds1.end(), 0.0);
            int s0 = 0;
            for (auto& tt : tls) s0 += tt.s0[j];
            for (auto& tt : tls)
                for (int64_t f = 0; f < p; ++f) ds1[f] += double(tt.s1[j * p + f]);
            cluster_s0[j] = s0;
            for (int64_t f = 0; f < p; ++f) cluster_s1[j * p + f] = F(ds1[f]);
        }
        // kmeansComputeCentroidsCandidates (kmeans_lloyd_impl.i:438-480)
        size_t cnum = 0;
        {
            std::vector<F> tv(k);
            std::vector<size_t> ti(k);
            for (auto& tt : tls) {
                size_t cpos = 0, lpos = 0;
                while (cpos + lpos < size_t(k) && (cpos < cnum || lpos < tt.cnum)) {
                    if (cpos < cnum && (lpos == tt.cnum || cvalues[cpos] > tt.cvalues[lpos])) {
                        tv[cpos + lpos] = cvalues[cpos];
                        ti[cpos + lpos] = cindices[cpos];
                        cpos++;
                    } else {
                        tv[cpos + lpos] = tt.cvalues[lpos];
                        ti[cpos + lpos] = tt.cindices[lpos];
                        lpos++;
                    }
                }
                cnum = cpos + lpos;
                std::copy(tv.begin(), tv.begin() + cnum, cvalues.begin());
                std::copy(ti.begin(), ti.begin() + cnum, cindices.begin());
            }
        }
        // update (kmeans_lloyd_batch_impl.i:142-188)
        size_t cpos = 0;
        F new_centers_goal(0);
        F l2norm(0);
        for (int64_t j = 0; j < k; ++j) {
            if (cluster_s0[j] > 0) {
                const F coeff = F(1.0 / cluster_s0[j]);
                for (int64_t f = 0; f < p; ++f) {
                    const F nc = cluster_s1[j * p + f] * coeff;
                    const F dd = in_clusters[j * p + f] - nc;
                    l2norm += dd * dd;
                    clusters[j * p + f] = nc;
                }
            } else {
                if (cpos >= cnum) {
                    status = 1;
                    break;
                }
                new_centers_goal += cvalues[cpos];
                const F* row = x + int64_t(cindices[cpos]) * p;
                for (int64_t f = 0; f < p; ++f) {
                    const F dd = in_clusters[j * p + f] - row[f];
                    l2norm += dd * dd;
                }
                std::memcpy(&clusters[j * p], row, sizeof(F) * p);
                cpos++;
            }
        }
        if (status) break;
        F new_target(0);
        for (auto& tt : tls) new_target += tt.goal;
        new_target -= new_centers_goal;
        in_clusters = clusters;
        if (stop_rule == 0) {
            // kmeans_lloyd_batch_impl.i:191-205
            if (l2norm <= F(thr)) {
                kiter++;
                break;
            }
            old_target = new_target;
        } else if (stop_rule == 1) {
            if (thr > 0 && double(new_target) + thr > double(prev_obj_gpu_rule)) {
                kiter++;
                break;
            }
            prev_obj_gpu_rule = new_target;
        }
    }
    (void)old_target;
    if (status) return status;

    // PostProcessing::computeAssignments (kmeans_lloyd_postprocessing.h:64-136)
    daal_cluster_halfsq(in_clusters.data(), k, p, clsq);
    std::vector<int32_t> assign(n);
    parallel_ranges(nblocks, nthr, [&](int, int64_t b0, int64_t b1) {
        std::vector<F> xc(size_t(block_size) * size_t(k));
        for (int64_t blk = b0; blk < b1; ++blk) {
            const int64_t row0 = blk * block_size;
            const int64_t bs = std::min(block_size, n - row0);
            const F* data = x + row0 * p;
            for (int64_t i = 0; i < bs; ++i)
                for (int64_t j = 0; j < k; ++j) {
                    F dot(0);
                    for (int64_t f = 0; f < p; ++f) dot += in_clusters[j * p + f] * data[i * p + f];
                    xc[i * k + j] = dot;
                }
            for (int64_t i = 0; i < bs; ++i) {
                F min_goal = clsq[0] - xc[i * k];
                int64_t min_idx = 0;
                for (int64_t j = 1; j < k; ++j) {
                    if (min_goal > clsq[j] - xc[i * k + j]) {
                        min_goal = clsq[j] - xc[i * k + j];
                        min_idx = j;
                    }
                }
                assign[row0 + i] = int32_t(min_idx);
            }
        }
    });
    // computeExactObjectiveFunction (kmeans_lloyd_postprocessing.h:138-183)
    std::vector<F> goal_local(nblocks);
    parallel_ranges(nblocks, nthr, [&](int, int64_t b0, int64_t b1) {
        for (int64_t blk = b0; blk < b1; ++blk) {
            const int64_t row0 = blk * block_size;
            const int64_t bs = std::min(block_size, n - row0);
            F goal(0);
            for (int64_t i = 0; i < bs; ++i) {
                const int64_t a = assign[row0 + i];
                for (int64_t f = 0; f < p; ++f) {
                    const F dd = x[(row0 + i) * p + f] - in_clusters[a * p + f];
                    goal += dd * dd;
                }
            }
            goal_local[blk] = goal;
        }
    });
    F obj(0);
    for (int64_t b = 0; b < nblocks; ++b) obj += goal_local[b];

    std::copy(in_clusters.begin(), in_clusters.end(), c_out);
    std::copy(assign.begin(), assign.end(), labels_out);
    *obj_out = double(obj);
    *iters_out = kiter;
    return 0;
}

// ---------------------------------------------------------------------------
// 2. canonical restatement of the MI355X engine's arithmetic
// ---------------------------------------------------------------------------
template <typename F>
inline F canon_dist(const F* x, const F* c, int64_t d) {
    F acc(0);
    for (int64_t f = 0; f < d; ++f) {
        const F diff = x[f] - c[f];
        const F sq = diff * diff;
        acc = acc + sq;
    }
    return acc;
}

template <typename F>
inline double dist64(const F* x, const F* c, int64_t d) {
    double acc = 0.0;
    for (int64_t f = 0; f < d; ++f) {
        const double diff = double(x[f]) - double(c[f]);
        const double sq = diff * diff;
        acc = acc + sq;
    }
    return acc;
}

template <typename F>
void canonical_assign(const F* x, int64_t n, int64_t d, const F* c, int64_t k, int32_t* labels,
                      F* best, int nthreads) {
    parallel_ranges(n, nthreads, [&](int, int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            F m = canon_dist(x + i * d, c, d);
            int32_t mi = 0;
            for (int64_t j = 1; j < k; ++j) {
                const F v = canon_dist(x + i * d, c + j * d, d);
                if (v < m) {
                    m = v;
                    mi = int32_t(j);
                }
            }
            labels[i] = mi;
            if (best) best[i] = m;
        }
    });
}

// objective in the engine's order: rows in chunks of obj_ch; partial t of a chunk
// sums rows t, t+256, ... in order; the 256 partials are combined by the tree
// p[t] += p[t+s] for s = 128, 64, ..., 1; chunk sums are added in chunk order.
template <typename F>
double canonical_objective(const F* x, int64_t n, int64_t d, const F* c, const int32_t* labels,
                           int64_t obj_ch, std::vector<double>* d64_out, int nthreads) {
    std::vector<double> d64(n);
    parallel_ranges(n, nthreads, [&](int, int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) d64[i] = dist64(x + i * d, c + int64_t(labels[i]) * d, d);
    });
    double total = 0.0;
    double part[256];
    for (int64_t r0 = 0; r0 < n; r0 += obj_ch) {
        const int64_t r1 = std::min(n, r0 + obj_ch);
        for (int t = 0; t < 256; ++t) {
            double p = 0.0;
            for (int64_t i = r0 + t; i < r1; i += 256) p = p + d64[i];
            part[t] = p;
        }
        for (int s = 128; s > 0; s >>= 1)
            for (int t = 0; t < s; ++t) part[t] = part[t] + part[t + s];
        total = total + part[0];
    }
    if (d64_out) *d64_out = std::move(d64);
    return total;
}

// centroid sums in the engine's order; returns counts
template <typename F>
void canonical_sums(const F* x, int64_t n, int64_t d, int64_t k, const int32_t* labels,
                    int64_t ch, int64_t g, std::vector<double>& s1, std::vector<int64_t>& counts) {
    std::vector<std::vector<int64_t>> members(k);
    counts.assign(k, 0);
    for (int64_t i = 0; i < n; ++i) counts[labels[i]]++;
    for (int64_t j = 0; j < k; ++j) members[j].reserve(counts[j]);
    for (int64_t i = 0; i < n; ++i) members[labels[i]].push_back(i);
    s1.assign(k * d, 0.0);
    std::vector<double> part(g * d), chunk(d);
    for (int64_t j = 0; j < k; ++j) {
        const auto& mem = members[j];
        const int64_t cnt = int64_t(mem.size());
        double* sj = s1.data() + j * d;
        for (int64_t q0 = 0; q0 < cnt; q0 += ch) {
            const int64_t q1 = std::min(cnt, q0 + ch);
            std::fill(part.begin(), part.end(), 0.0);
            for (int64_t m = q0; m < q1; ++m) {
                double* pg = part.data() + ((m - q0) % g) * d;
                const F* xr = x + mem[m] * d;
                for (int64_t f = 0; f < d; ++f) pg[f] = pg[f] + double(xr[f]);
            }
            for (int64_t f = 0; f < d; ++f) {
                double t = part[f];
                for (int64_t gg = 1; gg < g; ++gg) t = t + part[gg * d + f];
                chunk[f] = t;
            }
            for (int64_t f = 0; f < d; ++f) sj[f] = sj[f] + chunk[f];
        }
    }
}

template <typename F>
int canonical_lloyd(const F* x, int64_t n, int64_t d, const F* c0, int64_t k, int64_t max_iter,
                    double thr, int64_t ch, int64_t g, int64_t obj_ch, int nthreads, F* c_out,
                    int32_t* labels_out, double* obj_out, int64_t* iters_out,
                    double* iter_obj_out) {
    std::vector<F> c(c0, c0 + k * d);
    std::vector<int32_t> labels(n);
    std::vector<double> s1, d64;
    std::vector<int64_t> counts;
    double prev = double(std::numeric_limits<F>::max());
    int64_t iter = 0;
    for (iter = 0; iter < max_iter; ++iter) {
        canonical_assign(x, n, d, c.data(), k, labels.data(), (F*)nullptr, nthreads);
        double obj = canonical_objective(x, n, d, c.data(), labels.data(), obj_ch, &d64, nthreads);
        canonical_sums(x, n, d, k, labels.data(), ch, g, s1, counts);
        std::vector<F> cn(c);
        std::vector<int64_t> empties;
        for (int64_t j = 0; j < k; ++j) {
            if (counts[j] > 0) {
                for (int64_t f = 0; f < d; ++f) cn[j * d + f] = F(s1[j * d + f] / double(counts[j]));
            } else {
                empties.push_back(j);
            }
        }
        if (!empties.empty()) {
            const int64_t m = int64_t(empties.size());
            if (m > n) return 1;
            std::vector<int64_t> order(n);
            std::iota(order.begin(), order.end(), int64_t(0));
            std::partial_sort(order.begin(), order.begin() + m, order.end(),
                              [&](int64_t a, int64_t b) {
                                  if (d64[a] != d64[b]) return d64[a] > d64[b];
                                  return a < b;
                              });
            for (int64_t e = 0; e < m; ++e) {
                const int64_t j = empties[e];
                const int64_t src = order[e];
                std::memcpy(&cn[j * d], x + src * d, sizeof(F) * d);
                obj = obj - d64[src];
            }
        }
        c.swap(cn);
        if (iter_obj_out) iter_obj_out[iter] = obj;
        if (thr > 0 && obj + thr > prev) {
            iter++;
            break;
        }
        prev = obj;
    }
    canonical_assign(x, n, d, c.data(), k, labels.data(), (F*)nullptr, nthreads);
    const double obj = canonical_objective(x, n, d, c.data(), labels.data(), obj_ch,
                                           (std::vector<double>*)nullptr, nthreads);
    std::copy(c.begin(), c.end(), c_out);
    std::copy(labels.begin(), labels.end(), labels_out);
    *obj_out = obj;
    *iters_out = iter;
    return 0;
}

// ---------------------------------------------------------------------------
// 2b. canonical restatement for CSR rows (method lloyd_csr).  Follows the
//     reference's GPU CSR driver (cpp/oneapi/dal/algo/kmeans/backend/gpu/
//     train_kernel_lloyd_csr_dpc.cpp:131-226) and kernels (kernels_csr_impl.hpp):
//       - xs_i = sum of v*v over the row's stored entries in stored order (F)
//         (compute_data_squares, :38-55)
//       - cs_j = sum_f c_jf*c_jf in feature order (F) (compute_squares)
//       - s_ij = cs_j - 2*dot_ij, dot_ij = sum over stored entries in stored
//         order of v*c_j[col] (F, no FMA) (sparse gemm with alpha -2 + select,
//         :141-180); label = first j with the smallest s (strict '<' from j = 0:
//         ties go low, a NaN s_i0 keeps label 0, later NaNs never win)
//       - closest_i = s_i,label + xs_i (F) (complete_closest_distances, :182-186)
//       - objective = sum of double(closest_i) in the dense objective's order
//         (calc_objective_function, :193-203)
//       - centroid sums: the dense chunk/partial order of canonical_sums with the
//         stored entries of a member row added in stored order (update_centroids,
//         :219-322, adds them with atomics in no fixed order); counts 0 -> the
//         rows with the largest closest_i fill the empty clusters, densified as
//         copy_candidates_from_data does (empty_cluster_handling_dpc.cpp:152-189:
//         zeros, then the stored entries, a later duplicate overwriting)
//       - loop and stop rule of the dense driver above.
// ---------------------------------------------------------------------------
template <typename F>
struct csr_rows {
    const F* vals;
    const int64_t* cols;
    const int64_t* offs;
    int64_t base;
    int64_t begin(int64_t i) const { return offs[i] - base; }
    int64_t end(int64_t i) const { return offs[i + 1] - base; }
    int64_t col(int64_t p) const { return cols[p] - base; }
};

template <typename F>
void csr_assign(const csr_rows<F>& x, int64_t n, int64_t d, const F* c, int64_t k,
                int32_t* labels, F* closest, int nthreads) {
    std::vector<F> cs(k);
    for (int64_t j = 0; j < k; ++j) {
        F acc(0);
        for (int64_t f = 0; f < d; ++f) {
            const F sq = c[j * d + f] * c[j * d + f];
            acc = acc + sq;
        }
        cs[j] = acc;
    }
    parallel_ranges(n, nthreads, [&](int, int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            F xs(0);
            for (int64_t p = x.begin(i); p < x.end(i); ++p) {
                const F sq = x.vals[p] * x.vals[p];
                xs = xs + sq;
            }
            F best(0);
            int32_t lab = 0;
            for (int64_t j = 0; j < k; ++j) {
                F dot(0);
                for (int64_t p = x.begin(i); p < x.end(i); ++p) {
                    const F pr = x.vals[p] * c[j * d + x.col(p)];
                    dot = dot + pr;
                }
                const F two_dot = F(2) * dot;
                const F s = cs[j] - two_dot;
                if (j == 0 || s < best) {
                    if (j == 0) {
                        best = s;
                    } else {
                        best = s;
                        lab = int32_t(j);
                    }
                }
            }
            labels[i] = lab;
            closest[i] = best + xs;
        }
    });
}

// the per-row keys of the empty-cluster selection are closest_i with a negative
// (rounding) value or -0 taken as +0; the sum uses closest_i itself
template <typename F>
double csr_objective(const F* closest, int64_t n, int64_t obj_ch, std::vector<double>* d64_out) {
    std::vector<double> d64(n);
    for (int64_t i = 0; i < n; ++i) d64[i] = double(closest[i]);
    double total = 0.0;
    double part[256];
    for (int64_t r0 = 0; r0 < n; r0 += obj_ch) {
        const int64_t r1 = std::min(n, r0 + obj_ch);
        for (int t = 0; t < 256; ++t) {
            double p = 0.0;
            for (int64_t i = r0 + t; i < r1; i += 256) p = p + d64[i];
            part[t] = p;
        }
        for (int s = 128; s > 0; s >>= 1)
            for (int t = 0; t < s; ++t) part[t] = part[t] + part[t + s];
        total = total + part[0];
    }
    if (d64_out) {
        for (auto& v : d64) v = (v != v || v > 0.0) ? v : 0.0;
        *d64_out = std::move(d64);
    }
    return total;
}

template <typename F>
void csr_sums(const csr_rows<F>& x, int64_t n, int64_t d, int64_t k, const int32_t* labels,
              int64_t ch, int64_t g, std::vector<double>& s1, std::vector<int64_t>& counts) {
    std::vector<std::vector<int64_t>> members(k);
    counts.assign(k, 0);
    for (int64_t i = 0; i < n; ++i) counts[labels[i]]++;
    for (int64_t i = 0; i < n; ++i) members[labels[i]].push_back(i);
    s1.assign(k * d, 0.0);
    std::vector<double> part(g * d), chunk(d);
    for (int64_t j = 0; j < k; ++j) {
        const auto& mem = members[j];
        const int64_t cnt = int64_t(mem.size());
        double* sj = s1.data() + j * d;
        for (int64_t q0 = 0; q0 < cnt; q0 += ch) {
            const int64_t q1 = std::min(cnt, q0 + ch);
            std::fill(part.begin(), part.end(), 0.0);
            for (int64_t m = q0; m < q1; ++m) {
                double* pg = part.data() + ((m - q0) % g) * d;
                const int64_t row = mem[m];
                for (int64_t p = x.begin(row); p < x.end(row); ++p)
                    pg[x.col(p)] = pg[x.col(p)] + double(x.vals[p]);
            }
            for (int64_t f = 0; f < d; ++f) {
                double t = part[f];
                for (int64_t gg = 1; gg < g; ++gg) t = t + part[gg * d + f];
                chunk[f] = t;
            }
            for (int64_t f = 0; f < d; ++f) sj[f] = sj[f] + chunk[f];
        }
    }
}

template <typename F>
int csr_lloyd(const csr_rows<F>& x, int64_t n, int64_t d, const F* c0, int64_t k,
              int64_t max_iter, double thr, int64_t ch, int64_t g, int64_t obj_ch, int nthreads,
              F* c_out, int32_t* labels_out, double* obj_out, int64_t* iters_out) {
    std::vector<F> c(c0, c0 + k * d), closest(n);
    std::vector<int32_t> labels(n);
    std::vector<double> s1, d64;
    std::vector<int64_t> counts;
    double prev = double(std::numeric_limits<F>::max());
    int64_t iter = 0;
    for (iter = 0; iter < max_iter; ++iter) {
        csr_assign(x, n, d, c.data(), k, labels.data(), closest.data(), nthreads);
        double obj = csr_objective(closest.data(), n, obj_ch, &d64);
        csr_sums(x, n, d, k, labels.data(), ch, g, s1, counts);
        std::vector<F> cn(c);
        std::vector<int64_t> empties;
        for (int64_t j = 0; j < k; ++j) {
            if (counts[j] > 0) {
                for (int64_t f = 0; f < d; ++f) cn[j * d + f] = F(s1[j * d + f] / double(counts[j]));
            } else {
                empties.push_back(j);
            }
        }
        if (!empties.empty()) {
            const int64_t m = int64_t(empties.size());
            if (m > n) return 1;
            std::vector<int64_t> order(n);
            std::iota(order.begin(), order.end(), int64_t(0));
            // keys compare as the raw bits of the doubles (the engine's radix
            // select); for non-negative values that is the numeric order
            auto key = [&](int64_t i) {
                uint64_t b;
                std::memcpy(&b, &d64[i], sizeof(b));
                return b;
            };
            std::partial_sort(order.begin(), order.begin() + m, order.end(),
                              [&](int64_t a, int64_t b) {
                                  if (key(a) != key(b)) return key(a) > key(b);
                                  return a < b;
                              });
            for (int64_t e = 0; e < m; ++e) {
                const int64_t j = empties[e];
                const int64_t src = order[e];
                std::fill(cn.begin() + j * d, cn.begin() + (j + 1) * d, F(0));
                for (int64_t p = x.begin(src); p < x.end(src); ++p) cn[j * d + x.col(p)] = x.vals[p];
                obj = obj - d64[src];
            }
        }
        c.swap(cn);
        if (thr > 0 && obj + thr > prev) {
            iter++;
            break;
        }
        prev = obj;
    }
    csr_assign(x, n, d, c.data(), k, labels.data(), closest.data(), nthreads);
    const double obj = csr_objective(closest.data(), n, obj_ch, (std::vector<double>*)nullptr);
    std::copy(c.begin(), c.end(), c_out);
    std::copy(labels.begin(), labels.end(), labels_out);
    *obj_out = obj;
    *iters_out = iter;
    return 0;
}

// ---------------------------------------------------------------------------
// 3. FP64 brute force: labels and top-2 gap from exact-in-double distances
// ---------------------------------------------------------------------------
template <typename F>
void fp64_assign(const F* x, int64_t n, int64_t d, const F* c, int64_t k, int32_t* labels,
                 double* best, double* second, int nthreads) {
    parallel_ranges(n, nthreads, [&](int, int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            double m1 = std::numeric_limits<double>::infinity(), m2 = m1;
            int32_t i1 = 0;
            for (int64_t j = 0; j < k; ++j) {
                const double v = dist64(x + i * d, c + j * d, d);
                if (v < m1) {
                    m2 = m1;
                    m1 = v;
                    i1 = int32_t(j);
                } else if (v < m2) {
                    m2 = v;
                }
            }
            labels[i] = i1;
            if (best) best[i] = m1;
            if (second) second[i] = m2;
        }
    });
}

// ---------------------------------------------------------------------------
// 3. initial centroids: the engine's kmeans_init methods, restated (the reference
//    draws from DAAL's mt19937, backend/primitives/rng/host_engine.hpp:160-185,
//    so against the reference only the distribution can agree; the property the
//    reference's own test checks is that the centroids are distinct data rows,
//    kmeans_init/test/fixture.hpp:61-90)
// ---------------------------------------------------------------------------
// SplitMix64 finaliser (Steele, Lea, Flood, OOPSLA 2014), counter based:
// draw i of stream s under seed q = mix(mix(q ^ s * C1) ^ i * C2)
uint64_t init_mix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
uint64_t init_draw(uint64_t seed, uint64_t stream, uint64_t i) {
    return init_mix(init_mix(seed ^ (stream * 0xD1B54A32D192ED03ull)) ^ (i * 0x9E3779B97F4A7C15ull));
}
double init_unif(uint64_t seed, uint64_t stream, uint64_t i) {
    return double(init_draw(seed, stream, i) >> 11) / 9007199254740992.0;  // top 53 bits
}
uint64_t init_index(uint64_t seed, uint64_t stream, uint64_t i, uint64_t m) {  // in [0, m)
    return uint64_t((static_cast<unsigned __int128>(init_draw(seed, stream, i)) * m) >> 64);
}

// method 0: first k rows; 1: partial Fisher-Yates over [0, n) (stream 2, draw i
// picks position i + index(n - i)); 2: k-means++ (stream 1): first row index(n),
// then for t = 1..k-1 D^2 sampling with u = unif(t) over
//   D2_r = min over chosen c of sum_f ((double)x_f - (double)c_f)^2 (in order),
//   chunk sums of 4096 rows (thread t of 256 adds rows t, t+256, ... in order,
//   then the tree 128, 64, ..., 1), group sums of 64 chunk sums (in order),
//   total = group sums in order; one running sum descends: the first group whose
//   running sum exceeds u * total (else the last), in it the first such chunk
//   (else its last), in that chunk the first such block of 16 rows (block sums
//   in row order; else its last), in that block the first row with D2 > 0 whose
//   running sum exceeds u * total.  Fallbacks: the block's last row with D2 > 0,
//   else the chunk's, else the chunk's first row; total = 0: row
//   min(n - 1, floor(u * n)).
template <typename F>
void init_centroids(const F* x, int64_t n, int64_t d, int64_t k, int method, uint64_t seed, F* c) {
    auto copy_row = [&](int64_t src, int64_t dst) {
        std::copy(x + src * d, x + (src + 1) * d, c + dst * d);
    };
    if (method == 0) {
        for (int64_t i = 0; i < k; ++i) copy_row(i, i);
        return;
    }
    if (method == 1) {
        std::map<int64_t, int64_t> moved;
        auto at = [&](int64_t p) {
            auto it = moved.find(p);
            return it == moved.end() ? p : it->second;
        };
        for (int64_t i = 0; i < k; ++i) {
            const int64_t j = i + int64_t(init_index(seed, 2, uint64_t(i), uint64_t(n - i)));
            const int64_t vi = at(i), vj = at(j);
            copy_row(vj, i);
            moved[j] = vi;
        }
        return;
    }
    copy_row(int64_t(init_index(seed, 1, 0, uint64_t(n))), 0);
    std::vector<double> D2(size_t(n), 0.0);
    const int64_t nchunks = (n + 4095) / 4096;
    std::vector<double> chunk(static_cast<size_t>(nchunks));
    for (int64_t t = 1; t < k; ++t) {
        const F* ct = c + (t - 1) * d;
        for (int64_t r = 0; r < n; ++r) {
            double acc = 0.0;
            for (int64_t f = 0; f < d; ++f) {
                const double diff = double(x[r * d + f]) - double(ct[f]);
                const double sq = diff * diff;
                acc = acc + sq;
            }
            D2[size_t(r)] = t == 1 ? acc : std::fmin(D2[size_t(r)], acc);
        }
        for (int64_t q = 0; q < nchunks; ++q) {
            const int64_t r0 = q * 4096, rows = std::min<int64_t>(4096, n - r0);
            double part[256];
            for (int th = 0; th < 256; ++th) {
                double p = 0.0;
                for (int64_t r = th; r < rows; r += 256) p = p + D2[size_t(r0 + r)];
                part[th] = p;
            }
            for (int st = 128; st > 0; st >>= 1)
                for (int th = 0; th < st; ++th) part[th] = part[th] + part[th + st];
            chunk[size_t(q)] = part[0];
        }
        const int64_t ngroups = (nchunks + 63) / 64;
        std::vector<double> gsum(static_cast<size_t>(ngroups));
        for (int64_t g = 0; g < ngroups; ++g) {
            double sg = 0.0;
            for (int64_t q = g * 64; q < std::min<int64_t>(nchunks, g * 64 + 64); ++q) sg = sg + chunk[size_t(q)];
            gsum[size_t(g)] = sg;
        }
        double total = 0.0;
        for (int64_t g = 0; g < ngroups; ++g) total = total + gsum[size_t(g)];
        const double u = init_unif(seed, 1, uint64_t(t));
        int64_t idx;
        if (!(total > 0.0)) {
            const double pos = u * double(n);
            idx = pos >= double(n - 1) ? n - 1 : int64_t(pos);
        } else {
            const double target = u * total;
            double cum = 0.0;
            int64_t g = 0;
            for (; g + 1 < ngroups; ++g) {
                if (cum + gsum[size_t(g)] > target) break;
                cum = cum + gsum[size_t(g)];
            }
            const int64_t q1 = std::min<int64_t>(nchunks, g * 64 + 64);
            int64_t q = g * 64;
            for (; q + 1 < q1; ++q) {
                if (cum + chunk[size_t(q)] > target) break;
                cum = cum + chunk[size_t(q)];
            }
            const int64_t r0 = q * 4096, rows = std::min<int64_t>(4096, n - r0);
            const int64_t nb = (rows + 15) / 16;
            int64_t j = 0;
            for (; j + 1 < nb; ++j) {
                double sb = 0.0;
                for (int64_t r = j * 16; r < std::min<int64_t>(rows, j * 16 + 16); ++r) sb = sb + D2[size_t(r0 + r)];
                if (cum + sb > target) break;
                cum = cum + sb;
            }
            idx = -1;
            int64_t last_pos = -1;
            double sacc = cum;
            for (int64_t r = j * 16; r < std::min<int64_t>(rows, j * 16 + 16); ++r) {
                const double w = D2[size_t(r0 + r)];
                if (w > 0.0) last_pos = r0 + r;
                sacc = sacc + w;
                if (w > 0.0 && sacc > target) {
                    idx = r0 + r;
                    break;
                }
            }
            if (idx < 0 && last_pos < 0)
                for (int64_t r = rows - 1; r >= 0; --r)
                    if (D2[size_t(r0 + r)] > 0.0) {
                        last_pos = r0 + r;
                        break;
                    }
            if (idx < 0) idx = last_pos >= 0 ? last_pos : r0;
        }
        copy_row(idx, t);
    }
}

}  // namespace

extern "C" {

uint64_t oracle_init_mix(uint64_t z) { return init_mix(z); }

void oracle_init_centroids_f32(const float* x, int64_t n, int64_t d, int64_t k, int method,
                               uint64_t seed, float* c) {
    init_centroids<float>(x, n, d, k, method, seed, c);
}

void oracle_init_centroids_f64(const double* x, int64_t n, int64_t d, int64_t k, int method,
                               uint64_t seed, double* c) {
    init_centroids<double>(x, n, d, k, method, seed, c);
}

int oracle_daal_lloyd_f32(const float* x, int64_t n, int64_t d, const float* c0, int64_t k,
                          int64_t max_iter, double thr, int stop_rule, int64_t block_size,
                          int nthreads, float* c_out, int32_t* labels_out, double* obj_out,
                          int64_t* iters_out) {
    return daal_lloyd<float>(x, n, d, c0, k, max_iter, thr, stop_rule, block_size, nthreads, c_out,
                             labels_out, obj_out, iters_out);
}

int oracle_daal_lloyd_f64(const double* x, int64_t n, int64_t d, const double* c0, int64_t k,
                          int64_t max_iter, double thr, int stop_rule, int64_t block_size,
                          int nthreads, double* c_out, int32_t* labels_out, double* obj_out,
                          int64_t* iters_out) {
    return daal_lloyd<double>(x, n, d, c0, k, max_iter, thr, stop_rule, block_size, nthreads,
                              c_out, labels_out, obj_out, iters_out);
}

int oracle_csr_lloyd_f32(const float* vals, const int64_t* cols, const int64_t* offs, int64_t base,
                         int64_t n, int64_t d, const float* c0, int64_t k, int64_t max_iter,
                         double thr, int64_t ch, int64_t g, int64_t obj_ch, int nthreads,
                         float* c_out, int32_t* labels_out, double* obj_out, int64_t* iters_out) {
    const csr_rows<float> x{vals, cols, offs, base};
    return csr_lloyd(x, n, d, c0, k, max_iter, thr, ch, g, obj_ch, nthreads, c_out, labels_out,
                     obj_out, iters_out);
}

int oracle_csr_lloyd_f64(const double* vals, const int64_t* cols, const int64_t* offs, int64_t base,
                         int64_t n, int64_t d, const double* c0, int64_t k, int64_t max_iter,
                         double thr, int64_t ch, int64_t g, int64_t obj_ch, int nthreads,
                         double* c_out, int32_t* labels_out, double* obj_out, int64_t* iters_out) {
    const csr_rows<double> x{vals, cols, offs, base};
    return csr_lloyd(x, n, d, c0, k, max_iter, thr, ch, g, obj_ch, nthreads, c_out, labels_out,
                     obj_out, iters_out);
}

void oracle_csr_assign_f32(const float* vals, const int64_t* cols, const int64_t* offs,
                           int64_t base, int64_t n, int64_t d, const float* c, int64_t k,
                           int nthreads, int32_t* labels, float* closest) {
    csr_assign(csr_rows<float>{vals, cols, offs, base}, n, d, c, k, labels, closest, nthreads);
}

void oracle_csr_assign_f64(const double* vals, const int64_t* cols, const int64_t* offs,
                           int64_t base, int64_t n, int64_t d, const double* c, int64_t k,
                           int nthreads, int32_t* labels, double* closest) {
    csr_assign(csr_rows<double>{vals, cols, offs, base}, n, d, c, k, labels, closest, nthreads);
}

double oracle_csr_objective_f32(const float* closest, int64_t n, int64_t obj_ch) {
    return csr_objective(closest, n, obj_ch, (std::vector<double>*)nullptr);
}

double oracle_csr_objective_f64(const double* closest, int64_t n, int64_t obj_ch) {
    return csr_objective(closest, n, obj_ch, (std::vector<double>*)nullptr);
}

int oracle_canonical_lloyd_f32(const float* x, int64_t n, int64_t d, const float* c0, int64_t k,
                               int64_t max_iter, double thr, int64_t ch, int64_t g,
                               int64_t obj_ch, int nthreads, float* c_out, int32_t* labels_out,
                               double* obj_out, int64_t* iters_out, double* iter_obj_out) {
    return canonical_lloyd<float>(x, n, d, c0, k, max_iter, thr, ch, g, obj_ch, nthreads, c_out,
                                  labels_out, obj_out, iters_out, iter_obj_out);
}

int oracle_canonical_lloyd_f64(const double* x, int64_t n, int64_t d, const double* c0,
                               int64_t k, int64_t max_iter, double thr, int64_t ch, int64_t g,
                               int64_t obj_ch, int nthreads, double* c_out, int32_t* labels_out,
                               double* obj_out, int64_t* iters_out, double* iter_obj_out) {
    return canonical_lloyd<double>(x, n, d, c0, k, max_iter, thr, ch, g, obj_ch, nthreads, c_out,
                                   labels_out, obj_out, iters_out, iter_obj_out);
}

void oracle_canonical_assign_f32(const float* x, int64_t n, int64_t d, const float* c, int64_t k,
                                 int32_t* labels, float* best, int nthreads) {
    canonical_assign<float>(x, n, d, c, k, labels, best, nthreads);
}

void oracle_canonical_assign_f64(const double* x, int64_t n, int64_t d, const double* c,
                                 int64_t k, int32_t* labels, double* best, int nthreads) {
    canonical_assign<double>(x, n, d, c, k, labels, best, nthreads);
}

double oracle_canonical_objective_f32(const float* x, int64_t n, int64_t d, const float* c,
                                      const int32_t* labels, int64_t obj_ch, int nthreads) {
    return canonical_objective<float>(x, n, d, c, labels, obj_ch, nullptr, nthreads);
}

double oracle_canonical_objective_f64(const double* x, int64_t n, int64_t d, const double* c,
                                      const int32_t* labels, int64_t obj_ch, int nthreads) {
    return canonical_objective<double>(x, n, d, c, labels, obj_ch, nullptr, nthreads);
}

void oracle_canonical_sums_f32(const float* x, int64_t n, int64_t d, int64_t k,
                               const int32_t* labels, int64_t ch, int64_t g, double* s1_out,
                               int64_t* counts_out) {
    std::vector<double> s1;
    std::vector<int64_t> counts;
    canonical_sums<float>(x, n, d, k, labels, ch, g, s1, counts);
    std::copy(s1.begin(), s1.end(), s1_out);
    std::copy(counts.begin(), counts.end(), counts_out);
}

void oracle_canonical_sums_f64(const double* x, int64_t n, int64_t d, int64_t k,
                               const int32_t* labels, int64_t ch, int64_t g, double* s1_out,
                               int64_t* counts_out) {
    std::vector<double> s1;
    std::vector<int64_t> counts;
    canonical_sums<double>(x, n, d, k, labels, ch, g, s1, counts);
    std::copy(s1.begin(), s1.end(), s1_out);
    std::copy(counts.begin(), counts.end(), counts_out);
}

void oracle_fp64_assign_f32(const float* x, int64_t n, int64_t d, const float* c, int64_t k,
                            int32_t* labels, double* best, double* second, int nthreads) {
    fp64_assign<float>(x, n, d, c, k, labels, best, second, nthreads);
}

void oracle_fp64_assign_f64(const double* x, int64_t n, int64_t d, const double* c, int64_t k,
                            int32_t* labels, double* best, double* second, int nthreads) {
    fp64_assign<double>(x, n, d, c, k, labels, best, second, nthreads);
}

}  // extern "C"
