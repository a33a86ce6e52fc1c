// Device checks of the oneapi::dal::kmeans drop-in surface: the reference's gold
// dataset (test/fixture.hpp:156-203), its empty-cluster fixture
// (fixture.hpp:205-219) and its smoke-data training with default initialisation
// and inference objective (fixture.hpp:221-252), through dal::train / dal::infer
// on a HIP queue.
// argv[1]: text file "n d k" then data, initial centroids, expected centroids,
// expected responses, expected objective (whitespace separated).
#include <cmath>
#include <cstdio>
#include <fstream>
#include <limits>
#include <vector>

#include "oneapi/dal/algo/kmeans.hpp"
#include "oneapi/dal/algo/kmeans_init.hpp"

namespace dal = oneapi::dal;
namespace km = oneapi::dal::kmeans;

static int failures = 0;
#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);          \
            failures++;                                                          \
        }                                                                        \
    } while (0)

template <typename F>
static void gold(const char* path, dal::hip::queue& q) {
    std::ifstream in(path);
    int64_t n, d, k;
    in >> n >> d >> k;
    std::vector<F> x(n * d), c0(k * d), ec(k * d);
    std::vector<int> er(n);
    double eobj;
    for (auto& v : x) in >> v;
    for (auto& v : c0) in >> v;
    for (auto& v : ec) in >> v;
    for (auto& v : er) in >> v;
    in >> eobj;
    const auto desc = km::descriptor<F>().set_cluster_count(k).set_max_iteration_count(100).set_accuracy_threshold(0.0);
    const auto res = dal::train(q, desc, dal::homogen_table::wrap(x.data(), n, d),
                                dal::homogen_table::wrap(c0.data(), k, d));
    const auto c = dal::row_accessor<const F>(res.get_model().get_centroids()).pull();
    const auto r = dal::row_accessor<const int>(res.get_responses()).pull();
    // match map (fixture.hpp:677-706) and tolerances (fixture.hpp:182-202)
    std::vector<int> mm(k);
    for (int64_t i = 0; i < k; ++i) {
        double best = 1e300;
        for (int64_t j = 0; j < k; ++j) {
            double s = 0;
            for (int64_t f = 0; f < d; ++f) s += (ec[j * d + f] - c[i * d + f]) * (ec[j * d + f] - c[i * d + f]);
            if (s < best) { best = s; mm[i] = int(j); }
        }
    }
    for (int64_t i = 0; i < k; ++i)
        for (int64_t f = 0; f < d; ++f) {
            const double l = ec[mm[i] * d + f], v = c[i * d + f];
            CHECK(std::fabs(l - v) / (std::fabs(l) + std::fabs(v) + 1e-300) < 1e-7);
        }
    for (int64_t i = 0; i < n; ++i) CHECK(er[i] == mm[r[i]]);
    CHECK(std::fabs(res.get_objective_function_value() - eobj) / eobj < 1e-7);
    CHECK(res.get_iteration_count() == 100);
    // inference with the objective
    const auto idesc = km::descriptor<F>(k).set_result_options(km::result_options::compute_exact_objective_function);
    const auto ir = dal::infer(q, idesc, res.get_model(), dal::homogen_table::wrap(x.data(), n, d));
    const auto ri = dal::row_accessor<const int>(ir.get_responses()).pull();
    for (int64_t i = 0; i < n; ++i) CHECK(ri[i] == r[i]);
    CHECK(ir.get_objective_function_value() == res.get_objective_function_value());
}

template <typename F>
static void empty_clusters(dal::hip::queue& q) {
    F data[] = {-10, -9.5, -9, -8.5, -8, -1, 1, 9, 9.5, 10};
    F init[] = {-10, -10, -10};
    const auto res = dal::train(q, km::descriptor<F>(3).set_max_iteration_count(1),
                                dal::homogen_table::wrap(data, 10, 1), dal::homogen_table::wrap(init, 3, 1));
    const auto c = dal::row_accessor<const F>(res.get_model().get_centroids()).pull();
    const double want[] = {-1.65, 10, 9.5};
    for (int j = 0; j < 3; ++j) CHECK(std::fabs(c[j] - want[j]) <= 1e-6 * std::fabs(want[j]));
    const auto r = dal::row_accessor<const int>(res.get_responses()).pull();
    const int wr[] = {0, 0, 0, 0, 0, 0, 0, 2, 2, 1};
    for (int i = 0; i < 10; ++i) CHECK(r[i] == wr[i]);
}

static void smoke_infer(dal::hip::queue& q) {
    // fixture.hpp:221-252: the model {(-1.5,-1.5),(1.5,1.5)} on the smoke data has objective 4
    float x[] = {1, 1, 2, 2, 1, 2, 2, 1, -1, -1, -1, -2, -2, -1, -2, -2};
    float c[] = {-1.5f, -1.5f, 1.5f, 1.5f};
    const auto m = km::model<>().set_centroids(dal::homogen_table::wrap(c, 2, 2));
    const auto desc = km::descriptor<float>(2).set_result_options(km::result_options::compute_exact_objective_function);
    const auto ir = dal::infer(q, desc, m, dal::homogen_table::wrap(x, 8, 2));
    CHECK(std::fabs(ir.get_objective_function_value() - 4.0) < 1e-5 * 4.0);
    const auto r = dal::row_accessor<const int>(ir.get_responses()).pull();
    const int wr[] = {1, 1, 1, 1, 0, 0, 0, 0};
    for (int i = 0; i < 8; ++i) CHECK(r[i] == wr[i]);
}

// fixture.hpp:221-252 (run by test/batch.cpp:86-91): train without initial
// centroids (default k-means++ initialisation), k=2, 4 iterations, threshold
// 0.001, on the smoke data; centroids {(-1.5,-1.5),(1.5,1.5)} at the reference's
// strict tolerance 5 eps iterations 100 and responses {1,1,1,1,0,0,0,0}, then
// inference with the trained model gives objective 4.  The initialisation's draws
// are SplitMix64, not DAAL's mt19937, so the cluster order may be swapped: the
// centroids and the partition are matched up to that permutation.
template <typename F>
static void smoke_train_default_init(dal::hip::queue& q) {
    const F data[] = {1.0, 1.0, 2.0, 2.0, 1.0, 2.0, 2.0, 1.0, -1.0, -1.0, -1.0, -2.0, -2.0, -1.0, -2.0, -2.0};
    const auto x = dal::homogen_table::wrap(data, 8, 2);
    const auto desc = km::descriptor<F>(2).set_max_iteration_count(4).set_accuracy_threshold(0.001);
    const auto res = dal::train(q, desc, x);
    const auto c = dal::row_accessor<const F>(res.get_model().get_centroids()).pull();
    const auto r = dal::row_accessor<const int>(res.get_responses()).pull();
    const int want_r[] = {1, 1, 1, 1, 0, 0, 0, 0};
    const double want_c[] = {-1.5, -1.5, 1.5, 1.5};
    const bool swapped = r[0] == 0;
    const double tol = 5.0 * std::numeric_limits<F>::epsilon() * double(res.get_iteration_count()) * 100.0;
    CHECK(res.get_iteration_count() >= 1 && res.get_iteration_count() <= 4);
    for (int j = 0; j < 2; ++j)
        for (int f = 0; f < 2; ++f) {
            const double want = want_c[2 * (swapped ? 1 - j : j) + f];
            CHECK(std::fabs(double(c[2 * j + f]) - want) <= tol * std::fabs(want));
        }
    for (int i = 0; i < 8; ++i) CHECK(r[i] == (swapped ? 1 - want_r[i] : want_r[i]));
    const auto idesc = km::descriptor<F>(2).set_result_options(km::result_options::compute_exact_objective_function);
    const auto ir = dal::infer(q, idesc, res.get_model(), x);
    CHECK(std::fabs(ir.get_objective_function_value() - 4.0) < 1e-5 * 4.0);
    const auto ri = dal::row_accessor<const int>(ir.get_responses()).pull();
    for (int i = 0; i < 8; ++i) CHECK(ri[i] == r[i]);
    // the deprecated label accessors (train_types.hpp:116-121, infer_types.hpp:94-99)
#pragma GCC diagnostic push
#pragma GCC diagnostic ignored "-Wdeprecated-declarations"
    km::train_result<> tr = res;
    tr.set_labels(ir.get_responses());
    CHECK(tr.get_labels().get_row_count() == 8);
    km::infer_result<> ir2 = ir;
    ir2.set_labels(res.get_responses());
    CHECK(dal::row_accessor<const int>(ir2.get_labels()).pull()[0] == r[0]);
#pragma GCC diagnostic pop
}

// kmeans_init through dal::compute: the reference's test data and check
// (kmeans_init/test/batch.cpp:40-56, fixture.hpp:61-90: k distinct data rows),
// dense = first rows, and k-means train without initial centroids equal to train
// from dal::compute<plus_plus_dense> with the same seed.
template <typename F, typename M>
static void init_fixture(dal::hip::queue& q, bool first_rows) {
    namespace ki = oneapi::dal::kmeans_init;
    const F data[] = {1.0, 1.0, 2.0, 2.0, 1.0, 2.0, 2.0, 1.0, -1.0, -1.0, -1.0, -2.0, -2.0, -1.0, -2.0, -2.0};
    const auto x = dal::homogen_table::wrap(data, 8, 2);
    for (int seed = 0; seed < 16; ++seed) {
        const auto r = dal::compute(q, ki::descriptor<F, M>(4).set_seed(seed), x);
        const auto c = dal::row_accessor<const F>(r.get_centroids()).pull();
        CHECK(r.get_centroids().get_row_count() == 4 && r.get_centroids().get_column_count() == 2);
        bool used[8] = {};
        for (int i = 0; i < 4; ++i) {
            int hit = -1;
            for (int j = 0; j < 8 && hit < 0; ++j)
                if (!used[j] && data[2 * j] == c[2 * i] && data[2 * j + 1] == c[2 * i + 1]) hit = j;
            CHECK(hit >= 0);
            if (hit >= 0) used[hit] = true;
            if (first_rows) CHECK(hit == i);
        }
    }
}

static void init_checks(dal::hip::queue& q) {
    namespace ki = oneapi::dal::kmeans_init;
    init_fixture<float, ki::method::dense>(q, true);
    init_fixture<double, ki::method::dense>(q, true);
    init_fixture<float, ki::method::random_dense>(q, false);
    init_fixture<double, ki::method::plus_plus_dense>(q, false);
    init_fixture<float, ki::method::plus_plus_dense>(q, false);
    float data[16] = {0};
    bool threw = false;
    try {
        dal::compute(q, ki::descriptor<float, ki::method::parallel_plus_dense>(3),
                     dal::homogen_table::wrap(data, 8, 2));
    } catch (const dal::unimplemented&) {
        threw = true;
    }
    CHECK(threw);
    // train without initial centroids == train from compute<plus_plus_dense>(seed 777)
    std::vector<float> x(3000 * 4);
    for (size_t i = 0; i < x.size(); ++i) x[i] = float((i * 2654435761u) % 1000) / 100.0f + float(i % 4 == 0 ? (i / 400) % 5 : 0);
    const auto xt = dal::homogen_table::wrap(x.data(), 3000, 4);
    const auto desc = km::descriptor<float>(6).set_max_iteration_count(8);
    const auto r1 = dal::train(q, desc, xt);
    const auto c0 = dal::compute(q, ki::descriptor<float, ki::method::plus_plus_dense>(6), xt).get_centroids();
    const auto r2 = dal::train(q, desc, xt, c0);
    const auto a = dal::row_accessor<const float>(r1.get_model().get_centroids()).pull();
    const auto b = dal::row_accessor<const float>(r2.get_model().get_centroids()).pull();
    for (int i = 0; i < 24; ++i) CHECK(a[i] == b[i]);
    CHECK(r1.get_objective_function_value() == r2.get_objective_function_value());
}

// Device tables (the reference's USM tables): data wrapped over device memory
// trains and infers without host copies, results are device tables, and they
// equal the host-table results bit for bit.
template <typename F>
static void device_tables(dal::hip::queue& q) {
    const int64_t n = 5000, d = 12, k = 9;
    std::vector<F> x(n * d);
    for (int64_t i = 0; i < n * d; ++i) x[i] = F(((i * 2654435761ull) % 2000) / 100.0) + F((i / d) % 7);
    std::vector<F> c0(x.begin(), x.begin() + k * d);
    void* dx = nullptr;
    void* dc = nullptr;
    CHECK(km_device_alloc(q.get_context(), int64_t(sizeof(F)) * n * d, &dx) == KM_OK);
    CHECK(km_device_alloc(q.get_context(), int64_t(sizeof(F)) * k * d, &dc) == KM_OK);
    CHECK(km_memcpy(q.get_context(), dx, x.data(), int64_t(sizeof(F)) * n * d, KM_COPY_HOST_TO_DEVICE) == KM_OK);
    CHECK(km_memcpy(q.get_context(), dc, c0.data(), int64_t(sizeof(F)) * k * d, KM_COPY_HOST_TO_DEVICE) == KM_OK);
    {
        const auto xd = dal::homogen_table::wrap(q, static_cast<const F*>(dx), n, d);
        const auto cd = dal::homogen_table::wrap(q, static_cast<const F*>(dc), k, d);
        CHECK(xd.is_device() && xd.get_device() == q.get_device());
        const auto desc = km::descriptor<F>(k).set_max_iteration_count(6);
        const auto hr = dal::train(q, desc, dal::homogen_table::wrap(x.data(), n, d),
                                   dal::homogen_table::wrap(c0.data(), k, d));
        const auto dr = dal::train(q, desc, xd, cd);
        CHECK(dr.get_model().get_centroids().is_device() && dr.get_responses().is_device());
        const auto hc = dal::row_accessor<const F>(hr.get_model().get_centroids()).pull();
        const auto dcen = dal::row_accessor<const F>(dr.get_model().get_centroids()).pull(q);
        for (int64_t i = 0; i < k * d; ++i) CHECK(hc[i] == dcen[i]);
        const auto hl = dal::row_accessor<const int>(hr.get_responses()).pull();
        const auto dl = dal::row_accessor<const int>(dr.get_responses()).pull(q, {100, 200});
        for (int64_t i = 0; i < 100; ++i) CHECK(hl[100 + i] == dl[i]);
        CHECK(hr.get_objective_function_value() == dr.get_objective_function_value());
        // a device table cannot be pulled without a queue
        bool threw = false;
        try {
            (void)dal::row_accessor<const F>(xd).pull();
        } catch (const dal::invalid_argument&) {
            threw = true;
        }
        CHECK(threw);
        // no initial centroids: device k-means++ equals the host-table path
        const auto hr2 = dal::train(q, desc, dal::homogen_table::wrap(x.data(), n, d));
        const auto dr2 = dal::train(q, desc, xd);
        CHECK(hr2.get_objective_function_value() == dr2.get_objective_function_value());
        // inference on device data with the device model, and with a host model
        const auto idesc = km::descriptor<F>(k).set_result_options(km::result_options::compute_exact_objective_function);
        const auto hi = dal::infer(q, idesc, hr.get_model(), dal::homogen_table::wrap(x.data(), n, d));
        const auto di = dal::infer(q, idesc, dr.get_model(), xd);
        const auto mi = dal::infer(q, idesc, hr.get_model(), xd);
        CHECK(di.get_responses().is_device());
        const auto hil = dal::row_accessor<const int>(hi.get_responses()).pull();
        const auto dil = dal::row_accessor<const int>(di.get_responses()).pull(q);
        const auto mil = dal::row_accessor<const int>(mi.get_responses()).pull(q);
        for (int64_t i = 0; i < n; ++i) CHECK(hil[i] == dil[i] && hil[i] == mil[i]);
        CHECK(hi.get_objective_function_value() == di.get_objective_function_value());
        // host data with the device model
        const auto xi = dal::infer(q, idesc, dr.get_model(), dal::homogen_table::wrap(x.data(), n, d));
        CHECK(xi.get_objective_function_value() == hi.get_objective_function_value());
    }
    CHECK(km_device_free(q.get_context(), dx) == KM_OK);
    CHECK(km_device_free(q.get_context(), dc) == KM_OK);
}

// method lloyd_csr on csr_table: the reference's CSR known answer
// (kmeans/test/batch.cpp:404-478: objective 100, responses and centroids as the
// dense run's), host and device CSR tables, and the default initialisation.
template <typename F>
static void csr_checks(dal::hip::queue& q) {
    const F vals[] = {F(10), F(10), F(20)};
    const F dense[] = {F(10), 0, 0, 0, F(10), 0, 0, 0, F(20)};
    const F init[] = {F(5), F(5), 0, 0, 0, F(20)};
    for (int one : {0, 1}) {
        const int64_t cols[] = {0 + one, 1 + one, 2 + one};
        const int64_t offs[] = {0 + one, 1 + one, 2 + one, 3 + one};
        const auto ix = one ? dal::sparse_indexing::one_based : dal::sparse_indexing::zero_based;
        const auto xc = dal::csr_table::wrap(vals, cols, offs, 3, 3, ix);
        CHECK(xc.get_non_zero_count() == 3 && xc.get_kind() == dal::csr_table::kind());
        const auto cdesc = km::descriptor<F, km::method::lloyd_csr>(2)
                               .set_max_iteration_count(1)
                               .set_accuracy_threshold(0.0)
                               .set_result_options(km::result_options::compute_exact_objective_function);
        const auto ddesc = km::descriptor<F, km::method::lloyd_dense>(2)
                               .set_max_iteration_count(1)
                               .set_accuracy_threshold(0.0);
        const auto cr = dal::train(q, cdesc, xc, dal::homogen_table::wrap(init, 2, 3));
        const auto dr = dal::train(q, ddesc, dal::homogen_table::wrap(dense, 3, 3),
                                   dal::homogen_table::wrap(init, 2, 3));
        CHECK(cr.get_objective_function_value() == 100.0);
        CHECK(dr.get_objective_function_value() == 100.0);
        const auto cl = dal::row_accessor<const int>(cr.get_responses()).pull(q);
        const auto dl = dal::row_accessor<const int>(dr.get_responses()).pull(q);
        for (int i = 0; i < 3; ++i) CHECK(cl[i] == dl[i]);
        CHECK(cl[0] == 0 && cl[1] == 0 && cl[2] == 1);
        const auto cc = dal::row_accessor<const F>(cr.get_model().get_centroids()).pull(q);
        const auto dc = dal::row_accessor<const F>(dr.get_model().get_centroids()).pull(q);
        for (int i = 0; i < 6; ++i) CHECK(cc[i] == dc[i]);
        const auto ir = dal::infer(q, cdesc, cr.get_model(), xc);
        CHECK(ir.get_objective_function_value() == 100.0);
        // a dense table with the CSR method, and the reverse, are rejected
        bool threw = false;
        try {
            dal::train(q, cdesc, dal::homogen_table::wrap(dense, 3, 3), dal::homogen_table::wrap(init, 2, 3));
        } catch (const dal::invalid_argument&) {
            threw = true;
        }
        CHECK(threw);
        threw = false;
        try {
            dal::train(q, ddesc, xc, dal::homogen_table::wrap(init, 2, 3));
        } catch (const dal::invalid_argument&) {
            threw = true;
        }
        CHECK(threw);
    }
    // a larger CSR table, on the host and on the device: same results
    const int64_t n = 4000, d = 40, k = 8;
    std::vector<F> v;
    std::vector<int64_t> ci, ro{0};
    for (int64_t i = 0; i < n; ++i) {
        const int64_t g = i % k;
        for (int64_t f = 0; f < d; ++f)
            if ((f + g) % 5 == 0) {
                v.push_back(F(g * 3 + 1) + F(((i * 7919 + f) % 100) / 100.0));
                ci.push_back(f);
            }
        ro.push_back(int64_t(v.size()));
    }
    const int64_t nnz = int64_t(v.size());
    void *dv = nullptr, *dci = nullptr, *dro = nullptr;
    CHECK(km_device_alloc(q.get_context(), int64_t(sizeof(F)) * nnz, &dv) == KM_OK);
    CHECK(km_device_alloc(q.get_context(), 8 * nnz, &dci) == KM_OK);
    CHECK(km_device_alloc(q.get_context(), 8 * (n + 1), &dro) == KM_OK);
    CHECK(km_memcpy(q.get_context(), dv, v.data(), int64_t(sizeof(F)) * nnz, KM_COPY_HOST_TO_DEVICE) == KM_OK);
    CHECK(km_memcpy(q.get_context(), dci, ci.data(), 8 * nnz, KM_COPY_HOST_TO_DEVICE) == KM_OK);
    CHECK(km_memcpy(q.get_context(), dro, ro.data(), 8 * (n + 1), KM_COPY_HOST_TO_DEVICE) == KM_OK);
    {
        const auto hx = dal::csr_table::wrap(v.data(), ci.data(), ro.data(), n, d, dal::sparse_indexing::zero_based);
        const auto dx = dal::csr_table::wrap(q, static_cast<const F*>(dv), static_cast<const int64_t*>(dci),
                                             static_cast<const int64_t*>(dro), n, d, dal::sparse_indexing::zero_based);
        CHECK(dx.is_device() && dx.get_non_zero_count() == nnz);
        const auto desc = km::descriptor<F, km::method::lloyd_csr>(k).set_max_iteration_count(5);
        const auto hr = dal::train(q, desc, hx);  // default initial centroids (k-means++)
        const auto dr = dal::train(q, desc, dx, hr.get_model().get_centroids());
        const auto hr2 = dal::train(q, desc, hx, hr.get_model().get_centroids());
        CHECK(dr.get_iteration_count() == hr2.get_iteration_count());
        CHECK(dr.get_objective_function_value() == hr2.get_objective_function_value());
        const auto dl = dal::row_accessor<const int>(dr.get_responses()).pull(q);
        const auto hl = dal::row_accessor<const int>(hr2.get_responses()).pull(q);
        bool same = true;
        for (int64_t i = 0; i < n; ++i) same = same && dl[i] == hl[i];
        CHECK(same);
        const auto ll = dal::row_accessor<const int>(hr.get_responses()).pull(q);
        bool in_range = true;
        for (int64_t i = 0; i < n; ++i) in_range = in_range && ll[i] >= 0 && ll[i] < k;
        CHECK(in_range);
        // the default start of a CSR training is compute<plus_plus_dense> (seed 777)
        // on the rows' dense expansion, for host and device tables alike
        namespace ki = oneapi::dal::kmeans_init;
        std::vector<F> xd(size_t(n * d), F(0));
        for (int64_t i = 0; i < n; ++i)
            for (int64_t p = ro[i]; p < ro[i + 1]; ++p) xd[size_t(i * d + ci[p])] += v[p];
        const auto c0 = dal::compute(q, ki::descriptor<F, ki::method::plus_plus_dense>(k),
                                     dal::homogen_table::wrap(xd.data(), n, d))
                            .get_centroids();
        const auto er = dal::train(q, desc, hx, c0);
        const auto dr0 = dal::train(q, desc, dx);
        CHECK(er.get_objective_function_value() == hr.get_objective_function_value());
        CHECK(dr0.get_objective_function_value() == hr.get_objective_function_value());
        const auto ec = dal::row_accessor<const F>(er.get_model().get_centroids()).pull(q);
        const auto hc = dal::row_accessor<const F>(hr.get_model().get_centroids()).pull(q);
        const auto d0c = dal::row_accessor<const F>(dr0.get_model().get_centroids()).pull(q);
        bool same_c = true;
        for (int64_t i = 0; i < k * d; ++i) same_c = same_c && ec[i] == hc[i] && d0c[i] == hc[i];
        CHECK(same_c);
    }
    CHECK(km_device_free(q.get_context(), dv) == KM_OK);
    CHECK(km_device_free(q.get_context(), dci) == KM_OK);
    CHECK(km_device_free(q.get_context(), dro) == KM_OK);
}

// kmeans/test/serialization.cpp:104-116: train, serialize, deserialize, and
// inference with the original and the restored model agree (host and device
// centroids)
template <typename F>
static void serialization_infer(dal::hip::queue& q) {
    const int64_t n = 3000, d = 10, k = 7;
    std::vector<F> x(n * d);
    for (int64_t i = 0; i < n * d; ++i) x[i] = F(((i * 7919ull) % 1000) / 50.0) + F((i / d) % 7) * F(4);
    const auto xt = dal::homogen_table::wrap(x.data(), n, d);
    const auto desc = km::descriptor<F>(k).set_max_iteration_count(8).set_result_options(
        km::result_options::compute_exact_objective_function);
    const auto tr = dal::train(q, desc, xt, dal::homogen_table::wrap(x.data(), k, d));
    void* dx = nullptr;
    CHECK(km_device_alloc(q.get_context(), int64_t(sizeof(F)) * n * d, &dx) == KM_OK);
    CHECK(km_memcpy(q.get_context(), dx, x.data(), int64_t(sizeof(F)) * n * d, KM_COPY_HOST_TO_DEVICE) == KM_OK);
    const auto dtr = dal::train(q, desc, dal::homogen_table::wrap(q, static_cast<const F*>(dx), n, d),
                                dal::homogen_table::wrap(x.data(), k, d));
    for (const auto* m : {&tr.get_model(), &dtr.get_model()}) {
        dal::detail::binary_output_archive out;
        dal::detail::serialize(*m, out);
        const auto bytes = out.to_array();
        km::model<> back;
        dal::detail::binary_input_archive in(bytes.data(), int64_t(bytes.size()));
        dal::detail::deserialize(back, in);
        CHECK(!back.get_centroids().is_device());
        const auto a = dal::infer(q, desc, *m, xt);
        const auto b = dal::infer(q, desc, back, xt);
        CHECK(a.get_objective_function_value() == b.get_objective_function_value());
        const auto al = dal::row_accessor<const int>(a.get_responses()).pull(q);
        const auto bl = dal::row_accessor<const int>(b.get_responses()).pull(q);
        bool same = true;
        for (int64_t i = 0; i < n; ++i) same = same && al[i] == bl[i];
        CHECK(same);
    }
    CHECK(km_device_free(q.get_context(), dx) == KM_OK);
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    dal::hip::queue q(0);
    serialization_infer<float>(q);
    serialization_infer<double>(q);
    csr_checks<float>(q);
    csr_checks<double>(q);
    device_tables<float>(q);
    device_tables<double>(q);
    init_checks(q);
    gold<float>(argv[1], q);
    gold<double>(argv[1], q);
    empty_clusters<float>(q);
    empty_clusters<double>(q);
    smoke_infer(q);
    smoke_train_default_init<float>(q);
    smoke_train_default_init<double>(q);
    dal::hip::queue qb(0);
    qb.set_bf16_refine(true);
    gold<float>(argv[1], qb);
    empty_clusters<float>(qb);
    // no queue: default device
    float data[] = {0, 0, 0.1f, 0, 5, 5, 5.1f, 5};
    float init[] = {0, 0, 5, 5};
    const auto res = dal::train(km::descriptor<float>(2).set_max_iteration_count(3),
                                dal::homogen_table::wrap(data, 4, 2), dal::homogen_table::wrap(init, 2, 2));
    CHECK(res.get_iteration_count() == 3);
    if (failures) {
        std::printf("%d failure(s)\n", failures);
        return 1;
    }
    std::printf("api device checks passed\n");
    return 0;
}
