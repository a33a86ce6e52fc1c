// oneapi/dal/algo/kmeans_init.hpp -- drop-in C++ surface of oneDAL's kmeans_init
// (initial centroids for k-means) over km_init_centroids_* of the MI355X engine.
//
// Mirrors the reference interface:
//   task::init, method::dense / random_dense / plus_plus_dense /
//   parallel_plus_dense      cpp/oneapi/dal/algo/kmeans_init/common.hpp:25-60
//   descriptor (cluster_count 2, seed 777, local_trials_count -1)
//                            common.hpp:120-235, common.cpp:20-70
//   compute_input / compute_result
//                            compute_types.hpp:30-110
//   dal::compute             cpp/oneapi/dal/compute.hpp
// with the setter validation of common.cpp:50-68 and the preconditions of
// detail/compute_ops.hpp:40-56.  dense takes the first k rows (as the
// reference); random_dense and plus_plus_dense draw from the engine's counter
// generator keyed by the seed, so they match the reference in distribution,
// not draw for draw (its engine is DAAL's mt19937).  plus_plus_dense runs on
// the device here (the reference's GPU kernel throws unimplemented for it);
// parallel_plus_dense throws unimplemented, as on the reference's GPU.
// local_trials_count > 1 (greedy k-means++) is not implemented.
#pragma once

#include <cstdint>
#include <type_traits>

#include "../../../onedal_amd/kmeans_hip.h"
#include "../exceptions.hpp"
#include "../table.hpp"
#include "kmeans.hpp"

namespace oneapi::dal {
namespace kmeans_init {

namespace task {
struct init {};
using by_default = init;
}  // namespace task

namespace method {
struct dense {};
struct random_dense {};
struct plus_plus_dense {};
struct parallel_plus_dense {};
using by_default = dense;
}  // namespace method

template <typename Float = float, typename Method = method::by_default,
          typename Task = task::by_default>
class descriptor {
    static_assert(std::is_same_v<Float, float> || std::is_same_v<Float, double>,
                  "kmeans_init supports float and double");
    static_assert(std::is_same_v<Task, task::init>, "kmeans_init has the init task only");

public:
    using float_t = Float;
    using method_t = Method;
    using task_t = Task;

    explicit descriptor(std::int64_t cluster_count = 2) { set_cluster_count(cluster_count); }

    std::int64_t get_cluster_count() const { return cluster_count_; }
    descriptor& set_cluster_count(std::int64_t value) {
        if (value <= 0) throw domain_error(oneapi::dal::detail::error_messages::cluster_count_leq_zero());
        cluster_count_ = value;
        return *this;
    }
    std::int64_t get_seed() const { return seed_; }
    descriptor& set_seed(std::int64_t value) {
        seed_ = value;
        return *this;
    }
    std::int64_t get_local_trials_count() const { return local_trials_count_; }
    descriptor& set_local_trials_count(std::int64_t value = -1) {
        // common.cpp:63-68 (the reference reuses the cluster-count message)
        if (value <= 0) throw domain_error(oneapi::dal::detail::error_messages::cluster_count_leq_zero());
        local_trials_count_ = value;
        return *this;
    }

private:
    std::int64_t cluster_count_ = 2;
    std::int64_t seed_ = 777;
    std::int64_t local_trials_count_ = -1;
};

template <typename Task = task::by_default>
class compute_input {
public:
    using task_t = Task;
    compute_input(const table& data) : data_(data) {}
    const table& get_data() const { return data_; }
    compute_input& set_data(const table& value) {
        data_ = value;
        return *this;
    }

private:
    table data_;
};

template <typename Task = task::by_default>
class compute_result {
public:
    using task_t = Task;
    const table& get_centroids() const { return centroids_; }
    compute_result& set_centroids(const table& value) {
        centroids_ = value;
        return *this;
    }

private:
    table centroids_;
};

namespace detail {

template <typename Method>
constexpr std::int32_t method_id() {
    if constexpr (std::is_same_v<Method, method::dense>) return KM_INIT_DENSE;
    else if constexpr (std::is_same_v<Method, method::random_dense>) return KM_INIT_RANDOM_DENSE;
    else if constexpr (std::is_same_v<Method, method::plus_plus_dense>) return KM_INIT_PLUS_PLUS_DENSE;
    else return KM_INIT_PARALLEL_PLUS_DENSE;
}

// compute_ops::check_preconditions (detail/compute_ops.hpp:40-56)
template <typename Float, typename Method>
void check_compute_preconditions(const descriptor<Float, Method>& desc, const compute_input<>& in) {
    using namespace oneapi::dal::detail::error_messages;
    const table& xt = in.get_data();
    if (!xt.has_data()) throw domain_error(input_data_is_empty());
    if (desc.get_cluster_count() >= xt.get_row_count())
        throw domain_error(cluster_count_exceeds_data_row_count());
}

template <typename Float, typename Method>
compute_result<> compute_on(hip::queue& q, const descriptor<Float, Method>& desc,
                            const compute_input<>& in) {
    check_compute_preconditions(desc, in);
    const table& xt = in.get_data();
    const std::int64_t n = xt.get_row_count(), d = xt.get_column_count();
    const std::int64_t k = desc.get_cluster_count();
    if (desc.get_local_trials_count() > 1)
        throw unimplemented("K-Means init++ with several local trials is not implemented");
    const auto x = row_accessor<const Float>(xt).pull();
    array<Float> c(k * d);
    const auto seed = static_cast<std::uint64_t>(desc.get_seed());
    km_status st;
    if constexpr (std::is_same_v<Float, float>)
        st = km_init_centroids_f32(q.get_context(), x.get_data(), n, d, k, method_id<Method>(), seed, 0u,
                                   c.get_mutable_data());
    else
        st = km_init_centroids_f64(q.get_context(), x.get_data(), n, d, k, method_id<Method>(), seed, 0u,
                                   c.get_mutable_data());
    hip::check(st, q.get_context());
    return compute_result<>().set_centroids(homogen_table::wrap(c, k, d));
}

}  // namespace detail
}  // namespace kmeans_init

// dal::compute (compute.hpp)
template <typename Float, typename Method>
kmeans_init::compute_result<> compute(hip::queue& q, const kmeans_init::descriptor<Float, Method>& desc,
                                      const table& data) {
    return kmeans_init::detail::compute_on(q, desc, kmeans_init::compute_input<>(data));
}

template <typename Float, typename Method>
kmeans_init::compute_result<> compute(const kmeans_init::descriptor<Float, Method>& desc,
                                      const table& data) {
    return compute(hip::default_queue(), desc, data);
}

}  // namespace oneapi::dal
