// Host-side checks of the oneapi::dal::kmeans drop-in surface (no device call):
// descriptor setters, preconditions, result-option behaviour and tables, with the
// reference's exception types and messages (test/badarg.cpp:114-218,
// common.cpp:82-115, train_types.cpp:141-154, infer_types.cpp:91-110).
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
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

template <typename E>
static void expect_throw(const std::function<void()>& fn, const char* msg, int line) {
    try {
        fn();
        std::printf("FAIL line %d: no exception\n", line);
        failures++;
    } catch (const E& e) {
        if (msg && std::string(e.what()) != msg) {
            std::printf("FAIL line %d: message '%s' != '%s'\n", line, e.what(), msg);
            failures++;
        }
    } catch (const std::exception& e) {
        std::printf("FAIL line %d: wrong exception type: %s\n", line, e.what());
        failures++;
    }
}
#define EXPECT_THROW(E, expr, msg) expect_throw<E>([&] { expr; }, msg, __LINE__)

// Expected archive bytes, built field by field from the reference's serializers:
// prologue magic (detail/archives.hpp:23,35-39), polymorphic member = bool flag as
// one byte + uint64 id (detail/serialization.hpp:379-389), kmeans model impl id
// 8010000000 and its centroid table (algo/kmeans/common.cpp:44-56,144-156),
// homogen table = metadata, data, row count, column count, layout
// (table/backend/homogen_table_impl.hpp:136-138), simple metadata = dtypes, ftypes
// (table/common.cpp:133-141), array = array_id, element type, byte count, bytes
// (detail/array_impl.cpp:26-37); ids from backend/serialization.hpp:155-194,
// type codes from common.hpp:97-109, enums from table/common.hpp:35-36.
struct bytes_builder {
    std::vector<unsigned char> b;
    template <typename T>
    bytes_builder& put(T v) {
        unsigned char raw[sizeof(T)];
        std::memcpy(raw, &v, sizeof(T));
        b.insert(b.end(), raw, raw + sizeof(T));
        return *this;
    }
    bytes_builder& raw(const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        b.insert(b.end(), c, c + n);
        return *this;
    }
};

static std::vector<unsigned char> expected_model_bytes(int32_t dtype_code, int64_t rows, int64_t cols,
                                                       const void* data, int64_t elem) {
    bytes_builder e;
    e.put<uint32_t>(0x4441414Fu);
    e.put<uint8_t>(1).put<uint64_t>(8010000000ull);  // model impl
    if (rows == 0) {
        e.put<uint8_t>(1).put<uint64_t>(2020000000ull);  // empty table
        return e.b;
    }
    e.put<uint8_t>(1).put<uint64_t>(2030000000ull);  // homogen table
    e.put<uint8_t>(1).put<uint64_t>(2010000000ull);  // simple metadata
    e.put<uint64_t>(1000000000ull).put<int32_t>(0x4).put<int64_t>(4 * cols);
    for (int64_t j = 0; j < cols; ++j) e.put<int32_t>(dtype_code);
    e.put<uint64_t>(1000000000ull).put<int32_t>(0x4).put<int64_t>(4 * cols);
    const int32_t ftype = (dtype_code == 0x100 || dtype_code == 0x200) ? 3 : 1;  // ratio / ordinal
    for (int64_t j = 0; j < cols; ++j) e.put<int32_t>(ftype);
    e.put<uint64_t>(1000000000ull).put<int32_t>(0x10).put<int64_t>(rows * cols * elem);
    e.raw(data, size_t(rows * cols * elem));
    e.put<int64_t>(rows).put<int64_t>(cols).put<int32_t>(1);  // row_major
    return e.b;
}

static void archive_format_checks() {
    const double cd[] = {1.5, -2.0, 3.25, 0.0, 7.0, -1e-300};
    const float cf[] = {0.5f, -1.25f, 8.0f, 3.0f};
    {
        km::model<> m;
        m.set_centroids(dal::homogen_table::wrap(cd, 2, 3));
        dal::detail::binary_output_archive out;
        dal::detail::serialize(m, out);
        CHECK(out.to_array() == expected_model_bytes(0x200, 2, 3, cd, 8));
    }
    {
        km::model<> m;
        m.set_centroids(dal::homogen_table::wrap(cf, 4, 1));
        dal::detail::binary_output_archive out;
        dal::detail::serialize(m, out);
        CHECK(out.to_array() == expected_model_bytes(0x100, 4, 1, cf, 4));
        // and a hand-built archive deserializes to those centroids
        const auto e = expected_model_bytes(0x100, 4, 1, cf, 4);
        km::model<> back;
        dal::detail::binary_input_archive in(e.data(), int64_t(e.size()));
        dal::detail::deserialize(back, in);
        CHECK(back.get_cluster_count() == 4);
        CHECK(back.get_centroids().get_data_type() == dal::data_type::float32);
        const auto v = dal::row_accessor<const float>(back.get_centroids()).pull();
        for (int i = 0; i < 4; ++i) CHECK(v[i] == cf[i]);
    }
    {
        km::model<> empty;
        dal::detail::binary_output_archive out;
        dal::detail::serialize(empty, out);
        CHECK(out.to_array() == expected_model_bytes(0x100, 0, 0, nullptr, 4));
    }
    // an archive that claims more centroid bytes than it holds is refused before
    // any allocation of that size
    {
        bytes_builder e;
        e.put<uint32_t>(0x4441414Fu).put<uint8_t>(1).put<uint64_t>(8010000000ull);
        e.put<uint8_t>(1).put<uint64_t>(2030000000ull).put<uint8_t>(1).put<uint64_t>(2010000000ull);
        e.put<uint64_t>(1000000000ull).put<int32_t>(0x4).put<int64_t>(4).put<int32_t>(0x100);
        e.put<uint64_t>(1000000000ull).put<int32_t>(0x4).put<int64_t>(4).put<int32_t>(3);
        e.put<uint64_t>(1000000000ull).put<int32_t>(0x10).put<int64_t>(int64_t(1) << 50);
        e.put<int64_t>(int64_t(1) << 48).put<int64_t>(1).put<int32_t>(1);
        km::model<> t;
        dal::detail::binary_input_archive in(e.b.data(), int64_t(e.b.size()));
        EXPECT_THROW(dal::invalid_argument, dal::detail::deserialize(t, in),
                     "Archive content does not match requested type");
    }
    // a homogen table whose column types differ from its data size is refused
    {
        auto e = expected_model_bytes(0x200, 2, 3, cd, 8);
        // first dtype entry (after magic 4, flag+id 9, flag+id 9, flag+id 9, array header 20)
        const size_t at = 4 + 9 + 9 + 9 + 20;
        const int32_t f32 = 0x100;
        std::memcpy(&e[at], &f32, 4);
        km::model<> t;
        dal::detail::binary_input_archive in(e.data(), int64_t(e.size()));
        EXPECT_THROW(dal::invalid_argument, dal::detail::deserialize(t, in), nullptr);
    }
}

// model serialization round trip (the reference's kmeans/test/serialization.cpp
// trains, serializes, deserializes and compares inference; the inference half
// runs in test_api_gpu.cpp)
static void serialization_checks() {
    const double cent[] = {1.5, -2.0, 3.25, 0.0, 7.0, -1e-300};
    km::model<> m;
    m.set_centroids(dal::homogen_table::wrap(cent, 2, 3));
    dal::detail::binary_output_archive out;
    dal::detail::serialize(m, out);
    const auto bytes = out.to_array();
    CHECK(out.get_size() == int64_t(bytes.size()));
    km::model<> back;
    dal::detail::binary_input_archive in(bytes.data(), int64_t(bytes.size()));
    dal::detail::deserialize(back, in);
    const auto& c = back.get_centroids();
    CHECK(c.get_row_count() == 2 && c.get_column_count() == 3);
    CHECK(c.get_data_type() == dal::data_type::float64);
    const auto v = dal::row_accessor<const double>(c).pull();
    for (int i = 0; i < 6; ++i) CHECK(std::memcmp(&v[i], &cent[i], sizeof(double)) == 0);
    CHECK(back.get_cluster_count() == 2);
    // a truncated archive, and one that holds something else, are refused
    bool threw = false;
    try {
        km::model<> t;
        dal::detail::binary_input_archive cut(bytes.data(), int64_t(bytes.size()) - 1);
        dal::detail::deserialize(t, cut);
    } catch (const dal::invalid_argument&) {
        threw = true;
    }
    CHECK(threw);
    auto other = bytes;
    other[0] ^= 1;
    threw = false;
    try {
        km::model<> t;
        dal::detail::binary_input_archive wrong(other.data(), int64_t(other.size()));
        dal::detail::deserialize(t, wrong);
    } catch (const dal::invalid_argument&) {
        threw = true;
    }
    CHECK(threw);
    // an empty model round-trips to an empty model
    km::model<> empty, empty_back;
    dal::detail::binary_output_archive eo;
    dal::detail::serialize(empty, eo);
    const auto eb = eo.to_array();
    dal::detail::binary_input_archive ei(eb.data(), int64_t(eb.size()));
    dal::detail::deserialize(empty_back, ei);
    CHECK(!empty_back.get_centroids().has_data());
}

int main() {
    using namespace dal::detail::error_messages;
    serialization_checks();
    archive_format_checks();
    // descriptor defaults and setters
    km::descriptor<float> d;
    CHECK(d.get_cluster_count() == 2);
    CHECK(d.get_max_iteration_count() == 100);
    CHECK(d.get_accuracy_threshold() == 0.0);
    CHECK(!bool(d.get_result_options()));
    EXPECT_THROW(dal::domain_error, d.set_cluster_count(0), cluster_count_leq_zero());
    EXPECT_THROW(dal::domain_error, d.set_cluster_count(-1), cluster_count_leq_zero());
    EXPECT_THROW(dal::domain_error, d.set_cluster_count(int64_t(1) << 40), cluster_count_gt_max_int32());
    EXPECT_THROW(dal::domain_error, d.set_max_iteration_count(-1), max_iteration_count_lt_zero());
    EXPECT_THROW(dal::domain_error, d.set_accuracy_threshold(-0.5), accuracy_threshold_lt_zero());
    EXPECT_THROW(dal::domain_error, d.set_result_options(km::result_option_id{}), empty_set_of_result_options());
    // std:: base types are catchable too
    EXPECT_THROW(std::domain_error, km::descriptor<double>(0), nullptr);
    d.set_cluster_count(3).set_max_iteration_count(5).set_accuracy_threshold(0.01);
    CHECK(d.get_cluster_count() == 3 && d.get_max_iteration_count() == 5);

    // train preconditions, in the reference's order (detail/train_ops.hpp:40-61)
    float x[20];
    for (int i = 0; i < 20; ++i) x[i] = float(i);
    float c3[6] = {0, 0, 1, 1, 2, 2};
    float c2[4] = {0, 0, 1, 1};
    float c3x3[9] = {0};
    const auto X = dal::homogen_table::wrap(x, 10, 2);
    EXPECT_THROW(dal::domain_error, dal::train(d, dal::table{}, dal::homogen_table::wrap(c3, 3, 2)),
                 input_data_is_empty());
    EXPECT_THROW(dal::invalid_argument,
                 dal::train(km::descriptor<float>(11), X, dal::homogen_table::wrap(c3, 3, 2)),
                 cluster_count_exceeds_data_row_count());
    EXPECT_THROW(dal::invalid_argument, dal::train(d, X, dal::homogen_table::wrap(c2, 2, 2)),
                 input_initial_centroids_rc_neq_desc_cluster_count());
    EXPECT_THROW(dal::invalid_argument, dal::train(d, X, dal::homogen_table::wrap(c3x3, 3, 3)),
                 input_initial_centroids_cc_neq_input_data_cc());

    // infer preconditions (detail/infer_ops.hpp:40-56)
    km::model<> empty_model;
    EXPECT_THROW(dal::domain_error, dal::infer(d, empty_model, dal::table{}), input_data_is_empty());
    EXPECT_THROW(dal::domain_error, dal::infer(d, empty_model, X), input_model_centroids_are_empty());
    const auto m2 = km::model<>().set_centroids(dal::homogen_table::wrap(c2, 2, 2));
    EXPECT_THROW(dal::invalid_argument, dal::infer(d, m2, X),
                 input_model_centroids_rc_neq_desc_cluster_count());
    const auto m33 = km::model<>().set_centroids(dal::homogen_table::wrap(c3x3, 3, 3));
    EXPECT_THROW(dal::invalid_argument, dal::infer(d, m33, X),
                 input_model_centroids_cc_neq_input_data_cc());

    // results
    km::train_result<> tr;
    EXPECT_THROW(dal::domain_error, tr.set_iteration_count(-1), iteration_count_lt_zero());
    EXPECT_THROW(dal::domain_error, tr.set_objective_function_value(-1.0), objective_function_value_lt_zero());
    km::infer_result<> ir;
    EXPECT_THROW(dal::domain_error, ir.get_objective_function_value(),
                 this_result_is_not_enabled_via_result_options());
    EXPECT_THROW(dal::domain_error, ir.set_objective_function_value(1.0),
                 this_result_is_not_enabled_via_result_options());
    ir.set_result_options(km::result_options::compute_exact_objective_function);
    ir.set_objective_function_value(4.0);
    CHECK(ir.get_objective_function_value() == 4.0);
    EXPECT_THROW(dal::domain_error, ir.set_objective_function_value(-1.0), objective_function_value_lt_zero());

    // tables and accessors, with conversion
    CHECK(X.has_data() && X.get_row_count() == 10 && X.get_column_count() == 2);
    const auto rows = dal::row_accessor<const double>(X).pull({2, 4});
    CHECK(rows.get_count() == 4 && rows[0] == 4.0 && rows[3] == 7.0);
    int32_t lab[3] = {2, 0, 1};
    const auto L = dal::homogen_table::wrap(lab, 3, 1);
    const auto lf = dal::row_accessor<const float>(L).pull();
    CHECK(lf[0] == 2.0f && lf[2] == 1.0f);
    EXPECT_THROW(dal::out_of_range, dal::row_accessor<const float>(L).pull({1, 5}), nullptr);
    // model shares its table on copy (pimpl semantics)
    auto m = km::model<>().set_centroids(dal::homogen_table::wrap(c3, 3, 2));
    const auto m_copy = m;
    CHECK(m_copy.get_cluster_count() == 3);

    // kmeans_init descriptor (kmeans_init/common.cpp:20-68): defaults and setters
    namespace ki = oneapi::dal::kmeans_init;
    ki::descriptor<float, ki::method::plus_plus_dense> kd;
    CHECK(kd.get_cluster_count() == 2 && kd.get_seed() == 777 && kd.get_local_trials_count() == -1);
    EXPECT_THROW(dal::domain_error, kd.set_cluster_count(0), cluster_count_leq_zero());
    EXPECT_THROW(dal::domain_error, kd.set_local_trials_count(0), cluster_count_leq_zero());
    EXPECT_THROW(dal::domain_error, ki::descriptor<double>(-3), cluster_count_leq_zero());
    kd.set_cluster_count(4).set_seed(-5).set_local_trials_count(1);
    CHECK(kd.get_cluster_count() == 4 && kd.get_seed() == -5 && kd.get_local_trials_count() == 1);
    // preconditions run before any device call (detail/compute_ops.hpp:40-56)
    EXPECT_THROW(dal::domain_error, dal::kmeans_init::detail::check_compute_preconditions(
                     ki::descriptor<float>(3), ki::compute_input<>(dal::table{})), input_data_is_empty());
    float few[6] = {1, 2, 3, 4, 5, 6};
    EXPECT_THROW(dal::domain_error, dal::kmeans_init::detail::check_compute_preconditions(
                     ki::descriptor<float>(3), ki::compute_input<>(dal::homogen_table::wrap(few, 3, 2))),
                 cluster_count_exceeds_data_row_count());
    dal::kmeans_init::detail::check_compute_preconditions(
        ki::descriptor<float>(2), ki::compute_input<>(dal::homogen_table::wrap(few, 3, 2)));

    if (failures) {
        std::printf("%d failure(s)\n", failures);
        return 1;
    }
    std::printf("api host checks passed\n");
    return 0;
}
