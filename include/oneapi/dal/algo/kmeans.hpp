// oneapi/dal/algo/kmeans.hpp -- drop-in C++ surface of oneDAL's k-means
// (dense Lloyd) over the MI355X engine (include/onedal_amd/kmeans_hip.h).
//
// Mirrors the reference interface:
//   descriptor<Float, Method, Task>   cpp/oneapi/dal/algo/kmeans/common.hpp:155-220
//   model<Task>                       common.hpp:224-258
//   train_input / train_result        train_types.hpp:41-161
//   infer_input / infer_result        infer_types.hpp:42-138
//   result_options                    common.hpp:61-84
//   dal::train / dal::infer           cpp/oneapi/dal/train.hpp:26-59, infer.hpp
// with the same setter validation (common.cpp:82-115), preconditions
// (detail/train_ops.hpp:40-61, detail/infer_ops.hpp:40-56), result checks
// (train_types.cpp:141-154, infer_types.cpp:91-110) and messages.
//
// Where the reference takes a sycl::queue, this surface takes a
// oneapi::dal::hip::queue (one HIP device + stream); the overloads without a
// queue use device 0.  Tables hold host data (homogen_table::wrap(pointer, ...))
// or device data (homogen_table::wrap(queue, device_pointer, ...), the
// reference's USM tables).  With device data train and infer keep every buffer
// on the device (KM_FLAG_DEVICE_POINTERS) and return device tables, read on the
// host with row_accessor<const T>(table).pull(queue).
// Extension (no reference counterpart): hip::queue::set_bf16_refine(true)
// selects BF16 matrix-core distances with exact float re-score for float
// descriptors; results are identical to the default path.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../onedal_amd/kmeans_hip.h"
#include "../exceptions.hpp"
#include "../archive.hpp"
#include "../table.hpp"

namespace oneapi::dal {

// ---------------------------------------------------------------------------
// device queue
// ---------------------------------------------------------------------------
namespace hip {

class queue {
public:
    explicit queue(int device = 0) : impl_(std::make_shared<impl>(device)) {}
    int get_device() const { return impl_->device; }
    km_ctx* get_context() const { return impl_->ctx; }
    // extension: BF16 matrix-core distances + exact re-score (float data only)
    queue& set_bf16_refine(bool on) {
        impl_->bf16 = on;
        return *this;
    }
    bool get_bf16_refine() const { return impl_->bf16; }

private:
    struct impl {
        int device;
        km_ctx* ctx = nullptr;
        bool bf16 = false;
        explicit impl(int dev) : device(dev) {
            if (km_context_create(dev, &ctx) != KM_OK)
                throw unsupported_device("No usable HIP device " + std::to_string(dev));
        }
        ~impl() {
            if (ctx) km_context_destroy(ctx);
        }
        impl(const impl&) = delete;
        impl& operator=(const impl&) = delete;
    };
    std::shared_ptr<impl> impl_;
};

inline queue& default_queue() {
    static queue q(0);
    return q;
}

// Throw the oneapi::dal exception matching a km_status.
inline void check(km_status st, km_ctx* ctx) {
    if (st == KM_OK) return;
    const std::string msg = ctx ? km_last_error(ctx) : "invalid context";
    switch (st) {
        case KM_DOMAIN_ERROR: throw domain_error(msg);
        case KM_INVALID_ARGUMENT: throw invalid_argument(msg);
        case KM_DEVICE_BAD_ALLOC: throw device_bad_alloc();
        case KM_HOST_BAD_ALLOC: throw host_bad_alloc();
        case KM_UNIMPLEMENTED: throw unimplemented(msg);
        default: throw internal_error(msg);
    }
}

}  // namespace hip

// ---------------------------------------------------------------------------
// result option ids (util/result_option_id.hpp)
// ---------------------------------------------------------------------------
class result_option_id_base {
public:
    using bitset_t = std::uint64_t;
    constexpr result_option_id_base() : mask_(0) {}
    constexpr explicit result_option_id_base(bitset_t m) : mask_(m) {}
    constexpr explicit operator bool() const { return mask_ > 0; }
    constexpr bitset_t get_mask() const { return mask_; }
    constexpr bool test(const result_option_id_base& f) const { return (mask_ & f.mask_) != 0; }
    static constexpr result_option_id_base make_by_index(std::int64_t i) {
        return result_option_id_base(bitset_t(1) << i);
    }

private:
    bitset_t mask_;
};

namespace kmeans {

namespace task {
struct clustering {};
using by_default = clustering;
}  // namespace task

namespace method {
struct lloyd_dense {};
struct lloyd_csr {};
using by_default = lloyd_dense;
}  // namespace method

class result_option_id : public result_option_id_base {
public:
    constexpr result_option_id() = default;
    constexpr explicit result_option_id(const result_option_id_base& b) : result_option_id_base(b) {}
};

inline result_option_id operator|(const result_option_id& a, const result_option_id& b) {
    return result_option_id(result_option_id_base(a.get_mask() | b.get_mask()));
}

namespace result_options {
// common.cpp:24-26: bit 0
const inline result_option_id compute_exact_objective_function =
    result_option_id(result_option_id_base::make_by_index(0));
}  // namespace result_options

namespace detail {

template <typename Task = task::by_default>
class descriptor_base {
    static_assert(std::is_same_v<Task, task::clustering>);

public:
    using float_t = float;
    using method_t = method::by_default;
    using task_t = Task;

    descriptor_base() : impl_(std::make_shared<impl>()) {}
    std::int64_t get_cluster_count() const { return impl_->cluster_count; }
    std::int64_t get_max_iteration_count() const { return impl_->max_iteration_count; }
    double get_accuracy_threshold() const { return impl_->accuracy_threshold; }
    result_option_id get_result_options() const { return impl_->result_options; }

protected:
    void set_cluster_count_impl(std::int64_t v) {
        using namespace oneapi::dal::detail::error_messages;
        if (v <= 0) throw domain_error(cluster_count_leq_zero());
        if (v > std::numeric_limits<std::int32_t>::max()) throw domain_error(cluster_count_gt_max_int32());
        impl_->cluster_count = v;
    }
    void set_max_iteration_count_impl(std::int64_t v) {
        if (v < 0) throw domain_error(oneapi::dal::detail::error_messages::max_iteration_count_lt_zero());
        impl_->max_iteration_count = v;
    }
    void set_accuracy_threshold_impl(double v) {
        if (v < 0.0) throw domain_error(oneapi::dal::detail::error_messages::accuracy_threshold_lt_zero());
        impl_->accuracy_threshold = v;
    }
    void set_result_options_impl(const result_option_id& v) {
        if (!bool(v)) throw domain_error(oneapi::dal::detail::error_messages::empty_set_of_result_options());
        impl_->result_options = v;
    }

private:
    // defaults of common.cpp:36-42
    struct impl {
        std::int64_t cluster_count = 2;
        std::int64_t max_iteration_count = 100;
        double accuracy_threshold = 0.0;
        result_option_id result_options{};
    };
    std::shared_ptr<impl> impl_;
};

}  // namespace detail

template <typename Float = float, typename Method = method::by_default,
          typename Task = task::by_default>
class descriptor : public detail::descriptor_base<Task> {
    static_assert(std::is_same_v<Float, float> || std::is_same_v<Float, double>);
    static_assert(std::is_same_v<Method, method::lloyd_dense> ||
                  std::is_same_v<Method, method::lloyd_csr>);
    using base_t = detail::descriptor_base<Task>;

public:
    using float_t = Float;
    using method_t = Method;
    using task_t = Task;

    explicit descriptor(std::int64_t cluster_count = 2) { set_cluster_count(cluster_count); }
    std::int64_t get_cluster_count() const { return base_t::get_cluster_count(); }
    auto& set_cluster_count(std::int64_t v) {
        base_t::set_cluster_count_impl(v);
        return *this;
    }
    std::int64_t get_max_iteration_count() const { return base_t::get_max_iteration_count(); }
    auto& set_max_iteration_count(std::int64_t v) {
        base_t::set_max_iteration_count_impl(v);
        return *this;
    }
    double get_accuracy_threshold() const { return base_t::get_accuracy_threshold(); }
    auto& set_accuracy_threshold(double v) {
        base_t::set_accuracy_threshold_impl(v);
        return *this;
    }
    result_option_id get_result_options() const { return base_t::get_result_options(); }
    auto& set_result_options(const result_option_id& v) {
        base_t::set_result_options_impl(v);
        return *this;
    }
};

template <typename Task = task::by_default>
class model {
public:
    using task_t = Task;
    model() : impl_(std::make_shared<impl>()) {}
    const table& get_centroids() const { return impl_->centroids; }
    auto& set_centroids(const table& v) {
        impl_->centroids = v;
        return *this;
    }
    std::int64_t get_cluster_count() const { return impl_->centroids.get_row_count(); }

    // model<Task>::serialize / deserialize (common.cpp:44-56, 144-156), in the
    // reference's archive format: the model implementation as a polymorphic member
    // (present flag, kmeans_clustering_model_impl_id), whose one field is the
    // centroid table (an empty table, or a row-major homogeneous table).  Device
    // centroids are copied to the host first; deserialize yields host centroids.
    static constexpr std::int64_t serialization_id =
        std::int64_t(oneapi::dal::detail::serialization_ids::kmeans_clustering_model_impl_id);
    template <typename Archive>
    void serialize(Archive& ar) const {
        namespace dd = oneapi::dal::detail;
        const table& c = impl_->centroids;
        ar(std::uint8_t(1));
        ar(dd::serialization_ids::kmeans_clustering_model_impl_id);
        if (!c.has_data()) {
            dd::write_table(ar, 0, 0, 0, 0, nullptr);
            return;
        }
        const std::int64_t rows = c.get_row_count(), cols = c.get_column_count();
        const std::int64_t elem = std::int64_t(dd::size_of(c.get_data_type()));
        const std::int32_t code = c.get_data_type() == data_type::float32   ? dd::archived_type::float32
                                  : c.get_data_type() == data_type::float64 ? dd::archived_type::float64
                                                                            : dd::archived_type::int32;
        const std::int64_t bytes = rows * cols * elem;
        if (!c.is_device()) {
            dd::write_table(ar, rows, cols, code, elem, c.get_raw_data());
            return;
        }
        std::vector<unsigned char> host(static_cast<std::size_t>(bytes));
        km_ctx* ctx = nullptr;
        if (km_context_create(c.get_device(), &ctx) != KM_OK)
            throw internal_error("no HIP context for the device holding the centroids");
        const km_status st = km_memcpy(ctx, host.data(), c.get_raw_data(), bytes, KM_COPY_DEVICE_TO_HOST);
        km_context_destroy(ctx);
        if (st != KM_OK) throw internal_error("device copy of the centroids failed");
        dd::write_table(ar, rows, cols, code, elem, host.data());
    }
    template <typename Archive>
    void deserialize(Archive& ar) {
        namespace dd = oneapi::dal::detail;
        if (ar.template pop<std::uint8_t>() != 1 ||
            ar.template pop<std::uint64_t>() != dd::serialization_ids::kmeans_clustering_model_impl_id)
            throw invalid_argument(dd::archive_content_does_not_match_type());
        dd::archived_table t = dd::read_table(ar);
        table c;
        if (t.dtype_code == dd::archived_type::float32) c = to_table<float>(t);
        else if (t.dtype_code == dd::archived_type::float64) c = to_table<double>(t);
        else if (t.dtype_code == dd::archived_type::int32) c = to_table<std::int32_t>(t);
        impl_ = std::make_shared<impl>();
        impl_->centroids = c;
    }

private:
    template <typename T>
    static table to_table(const oneapi::dal::detail::archived_table& t) {
        array<T> a(t.rows * t.cols);
        std::memcpy(a.get_mutable_data(), t.bytes.data(), t.bytes.size());
        return homogen_table::wrap(a, t.rows, t.cols);
    }
    struct impl {
        table centroids;
    };
    std::shared_ptr<impl> impl_;
};

template <typename Task = task::by_default>
class train_input {
public:
    using task_t = Task;
    train_input(const table& data) : data_(data) {}
    train_input(const table& data, const table& initial_centroids)
            : data_(data), initial_centroids_(initial_centroids) {}
    const table& get_data() const { return data_; }
    auto& set_data(const table& v) {
        data_ = v;
        return *this;
    }
    const table& get_initial_centroids() const { return initial_centroids_; }
    auto& set_initial_centroids(const table& v) {
        initial_centroids_ = v;
        return *this;
    }

private:
    table data_, initial_centroids_;
};

template <typename Task = task::by_default>
class train_result {
public:
    using task_t = Task;
    const model<Task>& get_model() const { return model_; }
    auto& set_model(const model<Task>& v) {
        model_ = v;
        return *this;
    }
    [[deprecated]] const table& get_labels() const { return get_responses(); }
    [[deprecated]] auto& set_labels(const table& v) { return set_responses(v); }
    const table& get_responses() const { return responses_; }
    auto& set_responses(const table& v) {
        responses_ = v;
        return *this;
    }
    std::int64_t get_iteration_count() const { return iteration_count_; }
    auto& set_iteration_count(std::int64_t v) {
        if (v < 0) throw domain_error(oneapi::dal::detail::error_messages::iteration_count_lt_zero());
        iteration_count_ = v;
        return *this;
    }
    double get_objective_function_value() const { return objective_; }
    auto& set_objective_function_value(double v) {
        if (v < 0.0)
            throw domain_error(oneapi::dal::detail::error_messages::objective_function_value_lt_zero());
        objective_ = v;
        return *this;
    }

private:
    model<Task> model_;
    table responses_;
    std::int64_t iteration_count_ = 0;
    double objective_ = 0.0;
};

template <typename Task = task::by_default>
class infer_input {
public:
    using task_t = Task;
    infer_input(const model<Task>& m, const table& data) : model_(m), data_(data) {}
    const model<Task>& get_model() const { return model_; }
    auto& set_model(const model<Task>& v) {
        model_ = v;
        return *this;
    }
    const table& get_data() const { return data_; }
    auto& set_data(const table& v) {
        data_ = v;
        return *this;
    }

private:
    model<Task> model_;
    table data_;
};

template <typename Task = task::by_default>
class infer_result {
public:
    using task_t = Task;
    [[deprecated]] const table& get_labels() const { return get_responses(); }
    [[deprecated]] auto& set_labels(const table& v) { return set_responses(v); }
    const table& get_responses() const { return responses_; }
    auto& set_responses(const table& v) {
        responses_ = v;
        return *this;
    }
    double get_objective_function_value() const {
        if (!options_.test(result_options::compute_exact_objective_function))
            throw domain_error(
                oneapi::dal::detail::error_messages::this_result_is_not_enabled_via_result_options());
        return objective_;
    }
    auto& set_objective_function_value(double v) {
        using namespace oneapi::dal::detail::error_messages;
        if (!options_.test(result_options::compute_exact_objective_function))
            throw domain_error(this_result_is_not_enabled_via_result_options());
        if (v < 0.0) throw domain_error(objective_function_value_lt_zero());
        objective_ = v;
        return *this;
    }
    const result_option_id& get_result_options() const { return options_; }
    auto& set_result_options(const result_option_id& v) {
        options_ = v;
        return *this;
    }

private:
    table responses_;
    double objective_ = 0.0;
    result_option_id options_{};
};

namespace detail {

// train_ops::check_preconditions (detail/train_ops.hpp:40-61)
template <typename Desc>
void check_train_preconditions(const Desc& desc, const train_input<>& in) {
    using namespace oneapi::dal::detail::error_messages;
    const table& x = in.get_data();
    if (!x.has_data()) throw domain_error(input_data_is_empty());
    if (x.get_row_count() > std::numeric_limits<std::int32_t>::max())
        throw domain_error(row_count_gt_max_int32());
    if (desc.get_cluster_count() > x.get_row_count())
        throw invalid_argument(cluster_count_exceeds_data_row_count());
    const table& c0 = in.get_initial_centroids();
    if (c0.has_data()) {
        if (c0.get_row_count() != desc.get_cluster_count())
            throw invalid_argument(input_initial_centroids_rc_neq_desc_cluster_count());
        if (c0.get_column_count() != x.get_column_count())
            throw invalid_argument(input_initial_centroids_cc_neq_input_data_cc());
    }
}

// infer_ops::check_preconditions (detail/infer_ops.hpp:40-56)
template <typename Desc>
void check_infer_preconditions(const Desc& desc, const infer_input<>& in) {
    using namespace oneapi::dal::detail::error_messages;
    if (!in.get_data().has_data()) throw domain_error(input_data_is_empty());
    const table& c = in.get_model().get_centroids();
    if (!c.has_data()) throw domain_error(input_model_centroids_are_empty());
    if (c.get_row_count() != desc.get_cluster_count())
        throw invalid_argument(input_model_centroids_rc_neq_desc_cluster_count());
    if (c.get_column_count() != in.get_data().get_column_count())
        throw invalid_argument(input_model_centroids_cc_neq_input_data_cc());
}

// typed entry points of the C ABI
template <typename F>
km_status abi_train(km_ctx* ctx, const F* X, std::int64_t n, std::int64_t d, const F* C0, std::int64_t k,
                    std::int64_t max_iter, double thr, std::uint32_t flags, F* C, std::int32_t* L, double* obj,
                    std::int64_t* iters) {
    if constexpr (std::is_same_v<F, float>)
        return km_lloyd_train_f32(ctx, X, n, d, C0, k, max_iter, thr, flags, C, L, obj, iters);
    else
        return km_lloyd_train_f64(ctx, X, n, d, C0, k, max_iter, thr, flags, C, L, obj, iters);
}

template <typename F>
km_status abi_infer(km_ctx* ctx, const F* X, std::int64_t n, std::int64_t d, const F* C, std::int64_t k,
                    std::uint32_t flags, std::int32_t* L, double* obj) {
    if constexpr (std::is_same_v<F, float>)
        return km_lloyd_infer_f32(ctx, X, n, d, C, k, flags, L, obj);
    else
        return km_lloyd_infer_f64(ctx, X, n, d, C, k, flags, L, obj);
}

// Default initial centroids when train gets none.  The reference draws them with
// DAAL's k-means++ on the host, seed 777 (train_kernel_lloyd_dense_dpc.cpp:44-74,
// detail/train_init_centroids.hpp:47-74); here k-means++ runs on the device with
// the same seed.
template <typename F>
km_status abi_init_plus_plus(km_ctx* ctx, const F* X, std::int64_t n, std::int64_t d, std::int64_t k,
                             std::uint32_t flags, F* C) {
    if constexpr (std::is_same_v<F, float>)
        return km_init_centroids_f32(ctx, X, n, d, k, KM_INIT_PLUS_PLUS_DENSE, 777u, flags, C);
    else
        return km_init_centroids_f64(ctx, X, n, d, k, KM_INIT_PLUS_PLUS_DENSE, 777u, flags, C);
}

// Device memory of `count` T on the queue's device, freed with the last owner
// (the owner keeps the queue's context alive).
template <typename T>
std::shared_ptr<T> device_array(const hip::queue& q, std::int64_t count) {
    void* p = nullptr;
    hip::check(km_device_alloc(q.get_context(), std::int64_t(sizeof(T)) * count, &p), q.get_context());
    const hip::queue keep = q;
    return std::shared_ptr<T>(static_cast<T*>(p), [keep](T* ptr) { km_device_free(keep.get_context(), ptr); });
}

// A table's rows as Float in device memory: device tables of that element type
// are used in place, other tables are uploaded.
template <typename Float>
std::shared_ptr<const Float> to_device(const hip::queue& q, const table& t) {
    if (t.is_device()) {
        if (t.get_device() != q.get_device())
            throw invalid_argument("The queue's device does not hold the table");
        if (t.get_data_type() != oneapi::dal::detail::data_type_of<Float>())
            throw invalid_argument("A device table's element type must be the descriptor's float type");
        return std::shared_ptr<const Float>(static_cast<const Float*>(t.get_raw_data()), [](const Float*) {});
    }
    const auto h = row_accessor<const Float>(t).pull();
    auto dptr = device_array<Float>(q, h.get_count());
    hip::check(km_memcpy(q.get_context(), dptr.get(), h.get_data(), std::int64_t(sizeof(Float)) * h.get_count(),
                         KM_COPY_HOST_TO_DEVICE),
               q.get_context());
    return dptr;
}

template <typename F>
km_status abi_train_csr(km_ctx* ctx, const F* v, const std::int64_t* ci, const std::int64_t* ro, std::int64_t n,
                        std::int64_t d, std::int64_t nnz, std::int32_t ix, const F* C0, std::int64_t k,
                        std::int64_t max_iter, double thr, std::uint32_t flags, F* C, std::int32_t* L, double* obj,
                        std::int64_t* iters) {
    if constexpr (std::is_same_v<F, float>)
        return km_lloyd_train_csr_f32(ctx, v, ci, ro, n, d, nnz, ix, C0, k, max_iter, thr, flags, C, L, obj, iters);
    else
        return km_lloyd_train_csr_f64(ctx, v, ci, ro, n, d, nnz, ix, C0, k, max_iter, thr, flags, C, L, obj, iters);
}

template <typename F>
km_status abi_infer_csr(km_ctx* ctx, const F* v, const std::int64_t* ci, const std::int64_t* ro, std::int64_t n,
                        std::int64_t d, std::int64_t nnz, std::int32_t ix, const F* C, std::int64_t k,
                        std::uint32_t flags, std::int32_t* L, double* obj) {
    if constexpr (std::is_same_v<F, float>)
        return km_lloyd_infer_csr_f32(ctx, v, ci, ro, n, d, nnz, ix, C, k, flags, L, obj);
    else
        return km_lloyd_infer_csr_f64(ctx, v, ci, ro, n, d, nnz, ix, C, k, flags, L, obj);
}

// The CSR arrays of a table as the engine takes them: device tables in place
// (values of the descriptor's float type only), host tables by pointer (the
// engine uploads them; values converted to Float if needed).
template <typename Float>
struct csr_args {
    const Float* values = nullptr;
    const std::int64_t* cols = nullptr;
    const std::int64_t* offs = nullptr;
    std::int64_t nnz = 0;
    std::int32_t indexing = KM_SPARSE_ONE_BASED;
    std::vector<Float> converted;
};

template <typename Float>
void csr_of(const hip::queue& q, const table& t, csr_args<Float>& a) {
    if (t.get_kind() != csr_table::kind())
        throw invalid_argument("Method lloyd_csr needs a csr_table");
    const csr_table c(t);
    a.cols = c.get_column_indices();
    a.offs = c.get_row_offsets();
    a.nnz = c.get_non_zero_count();
    a.indexing = c.get_indexing() == sparse_indexing::zero_based ? KM_SPARSE_ZERO_BASED : KM_SPARSE_ONE_BASED;
    if (c.get_data_type() == oneapi::dal::detail::data_type_of<Float>()) {
        a.values = static_cast<const Float*>(c.get_data());
        return;
    }
    if (c.is_device())
        throw invalid_argument("A device table's element type must be the descriptor's float type");
    a.converted.resize(std::size_t(a.nnz));
    if (c.get_data_type() == data_type::float32) {
        const float* s = static_cast<const float*>(c.get_data());
        for (std::int64_t i = 0; i < a.nnz; ++i) a.converted[i] = Float(s[i]);
    } else {
        const double* s = static_cast<const double*>(c.get_data());
        for (std::int64_t i = 0; i < a.nnz; ++i) a.converted[i] = Float(s[i]);
    }
    a.values = a.converted.data();
    (void)q;
}

// Default initial centroids of CSR data: the reference runs DAAL's plusPlusCSR
// on the host (train_kernel_lloyd_csr_dpc.cpp:41-68, train_init_centroids.hpp:40-41);
// here the device k-means++ walks the CSR rows (the dense initialiser's result on
// the rows' dense expansion, with its seed), host or device arrays as given.
template <typename F>
km_status abi_init_plus_plus_csr(km_ctx* ctx, const csr_args<F>& a, std::int64_t n, std::int64_t d,
                                 std::int64_t k, std::uint32_t flags, F* C) {
    if constexpr (std::is_same_v<F, float>)
        return km_init_centroids_csr_f32(ctx, a.values, a.cols, a.offs, n, d, a.nnz, a.indexing, k,
                                         KM_INIT_PLUS_PLUS_DENSE, 777u, flags, C);
    else
        return km_init_centroids_csr_f64(ctx, a.values, a.cols, a.offs, n, d, a.nnz, a.indexing, k,
                                         KM_INIT_PLUS_PLUS_DENSE, 777u, flags, C);
}

template <typename Float>
train_result<> train_csr_on(hip::queue& q, const descriptor<Float, method::lloyd_csr>& desc,
                            const train_input<>& in) {
    const table& xt = in.get_data();
    const std::int64_t n = xt.get_row_count(), d = xt.get_column_count();
    const std::int64_t k = desc.get_cluster_count();
    km_ctx* ctx = q.get_context();
    csr_args<Float> a;
    csr_of<Float>(q, xt, a);
    double obj = 0.0;
    std::int64_t iters = 0;
    train_result<> r;
    const bool has_c0 = in.get_initial_centroids().has_data();
    if (xt.is_device()) {
        const std::uint32_t flags = KM_FLAG_DEVICE_POINTERS;
        std::shared_ptr<const Float> c0;
        if (has_c0) {
            c0 = to_device<Float>(q, in.get_initial_centroids());
        } else {
            auto dc = device_array<Float>(q, k * d);
            hip::check(abi_init_plus_plus_csr<Float>(ctx, a, n, d, k, KM_FLAG_DEVICE_POINTERS, dc.get()), ctx);
            c0 = dc;
        }
        auto c = device_array<Float>(q, k * d);
        auto labels = device_array<std::int32_t>(q, n);
        hip::check(abi_train_csr<Float>(ctx, a.values, a.cols, a.offs, n, d, a.nnz, a.indexing, c0.get(), k,
                                        desc.get_max_iteration_count(), desc.get_accuracy_threshold(), flags,
                                        c.get(), labels.get(), &obj, &iters),
                   ctx);
        r.set_model(model<>().set_centroids(homogen_table::wrap_device<Float>(c, k, d, q.get_device())));
        r.set_responses(homogen_table::wrap_device<std::int32_t>(labels, n, 1, q.get_device()));
    } else {
        array<Float> c0;
        if (has_c0) {
            c0 = row_accessor<const Float>(in.get_initial_centroids()).pull(q);
        } else {
            c0 = array<Float>(k * d);
            hip::check(abi_init_plus_plus_csr<Float>(ctx, a, n, d, k, 0u, c0.get_mutable_data()), ctx);
        }
        array<Float> c(k * d);
        array<std::int32_t> labels(n);
        hip::check(abi_train_csr<Float>(ctx, a.values, a.cols, a.offs, n, d, a.nnz, a.indexing, c0.get_data(), k,
                                        desc.get_max_iteration_count(), desc.get_accuracy_threshold(), 0u,
                                        c.get_mutable_data(), labels.get_mutable_data(), &obj, &iters),
                   ctx);
        r.set_model(model<>().set_centroids(homogen_table::wrap(c, k, d)));
        r.set_responses(homogen_table::wrap(labels, n, 1));
    }
    r.set_iteration_count(iters);
    r.set_objective_function_value(obj);
    return r;
}

template <typename Float>
infer_result<> infer_csr_on(hip::queue& q, const descriptor<Float, method::lloyd_csr>& desc,
                            const infer_input<>& in) {
    const table& xt = in.get_data();
    const std::int64_t n = xt.get_row_count(), d = xt.get_column_count();
    const std::int64_t k = desc.get_cluster_count();
    km_ctx* ctx = q.get_context();
    csr_args<Float> a;
    csr_of<Float>(q, xt, a);
    const bool want_obj = desc.get_result_options().test(result_options::compute_exact_objective_function);
    std::uint32_t flags = want_obj ? KM_FLAG_EXACT_OBJECTIVE : 0u;
    double obj = 0.0;
    infer_result<> r;
    r.set_result_options(desc.get_result_options());
    if (xt.is_device()) {
        flags |= KM_FLAG_DEVICE_POINTERS;
        const auto c = to_device<Float>(q, in.get_model().get_centroids());
        auto labels = device_array<std::int32_t>(q, n);
        hip::check(abi_infer_csr<Float>(ctx, a.values, a.cols, a.offs, n, d, a.nnz, a.indexing, c.get(), k, flags,
                                        labels.get(), &obj),
                   ctx);
        r.set_responses(homogen_table::wrap_device<std::int32_t>(labels, n, 1, q.get_device()));
    } else {
        const auto c = row_accessor<const Float>(in.get_model().get_centroids()).pull(q);
        array<std::int32_t> labels(n);
        hip::check(abi_infer_csr<Float>(ctx, a.values, a.cols, a.offs, n, d, a.nnz, a.indexing, c.get_data(), k,
                                        flags, labels.get_mutable_data(), &obj),
                   ctx);
        r.set_responses(homogen_table::wrap(labels, n, 1));
    }
    if (want_obj) r.set_objective_function_value(obj);
    return r;
}

template <typename Float, typename Method>
train_result<> train_dense_on(hip::queue& q, const descriptor<Float, Method>& desc,
                              const train_input<>& in) {
    if (in.get_data().get_kind() != homogen_table::kind())
        throw invalid_argument("Method lloyd_dense needs a homogeneous table");
    const table& xt = in.get_data();
    const std::int64_t n = xt.get_row_count(), d = xt.get_column_count();
    const std::int64_t k = desc.get_cluster_count();
    km_ctx* ctx = q.get_context();
    const bool has_c0 = in.get_initial_centroids().has_data();
    double obj = 0.0;
    std::int64_t iters = 0;
    train_result<> r;
    if (xt.is_device()) {
        // device data: every buffer stays on the device, results are device tables
        const std::uint32_t flags = KM_FLAG_DEVICE_POINTERS | (q.get_bf16_refine() ? KM_FLAG_BF16_REFINE : 0u);
        const auto x = to_device<Float>(q, xt);
        std::shared_ptr<const Float> c0;
        if (has_c0) {
            c0 = to_device<Float>(q, in.get_initial_centroids());
        } else {
            auto ci = device_array<Float>(q, k * d);
            hip::check(abi_init_plus_plus<Float>(ctx, x.get(), n, d, k, KM_FLAG_DEVICE_POINTERS, ci.get()), ctx);
            c0 = ci;
        }
        auto c = device_array<Float>(q, k * d);
        auto labels = device_array<std::int32_t>(q, n);
        hip::check(abi_train<Float>(ctx, x.get(), n, d, c0.get(), k, desc.get_max_iteration_count(),
                                    desc.get_accuracy_threshold(), flags, c.get(), labels.get(), &obj, &iters),
                   ctx);
        r.set_model(model<>().set_centroids(homogen_table::wrap_device<Float>(c, k, d, q.get_device())));
        r.set_responses(homogen_table::wrap_device<std::int32_t>(labels, n, 1, q.get_device()));
    } else {
        const std::uint32_t flags = q.get_bf16_refine() ? KM_FLAG_BF16_REFINE : 0u;
        const auto x = row_accessor<const Float>(xt).pull();
        array<Float> c0;
        if (has_c0) {
            c0 = row_accessor<const Float>(in.get_initial_centroids()).pull(q);
        } else {
            c0 = array<Float>(k * d);
            hip::check(abi_init_plus_plus<Float>(ctx, x.get_data(), n, d, k, 0u, c0.get_mutable_data()), ctx);
        }
        array<Float> c(k * d);
        array<std::int32_t> labels(n);
        hip::check(abi_train<Float>(ctx, x.get_data(), n, d, c0.get_data(), k, desc.get_max_iteration_count(),
                                    desc.get_accuracy_threshold(), flags, c.get_mutable_data(),
                                    labels.get_mutable_data(), &obj, &iters),
                   ctx);
        r.set_model(model<>().set_centroids(homogen_table::wrap(c, k, d)));
        r.set_responses(homogen_table::wrap(labels, n, 1));
    }
    r.set_iteration_count(iters);
    r.set_objective_function_value(obj);
    return r;
}

template <typename Float, typename Method>
infer_result<> infer_dense_on(hip::queue& q, const descriptor<Float, Method>& desc,
                              const infer_input<>& in) {
    if (in.get_data().get_kind() != homogen_table::kind())
        throw invalid_argument("Method lloyd_dense needs a homogeneous table");
    const table& xt = in.get_data();
    const std::int64_t n = xt.get_row_count(), d = xt.get_column_count();
    const std::int64_t k = desc.get_cluster_count();
    km_ctx* ctx = q.get_context();
    const bool want_obj = desc.get_result_options().test(result_options::compute_exact_objective_function);
    std::uint32_t flags = (want_obj ? KM_FLAG_EXACT_OBJECTIVE : 0u) | (q.get_bf16_refine() ? KM_FLAG_BF16_REFINE : 0u);
    double obj = 0.0;
    infer_result<> r;
    r.set_result_options(desc.get_result_options());
    if (xt.is_device()) {
        flags |= KM_FLAG_DEVICE_POINTERS;
        const auto x = to_device<Float>(q, xt);
        const auto c = to_device<Float>(q, in.get_model().get_centroids());
        auto labels = device_array<std::int32_t>(q, n);
        hip::check(abi_infer<Float>(ctx, x.get(), n, d, c.get(), k, flags, labels.get(), &obj), ctx);
        r.set_responses(homogen_table::wrap_device<std::int32_t>(labels, n, 1, q.get_device()));
    } else {
        const auto x = row_accessor<const Float>(xt).pull();
        const auto c = row_accessor<const Float>(in.get_model().get_centroids()).pull(q);
        array<std::int32_t> labels(n);
        hip::check(abi_infer<Float>(ctx, x.get_data(), n, d, c.get_data(), k, flags, labels.get_mutable_data(), &obj),
                   ctx);
        r.set_responses(homogen_table::wrap(labels, n, 1));
    }
    if (want_obj) r.set_objective_function_value(obj);
    return r;
}

template <typename Float, typename Method>
train_result<> train_on(hip::queue& q, const descriptor<Float, Method>& desc,
                        const train_input<>& in) {
    check_train_preconditions(desc, in);
    if constexpr (std::is_same_v<Method, method::lloyd_csr>)
        return train_csr_on<Float>(q, desc, in);
    else
        return train_dense_on<Float>(q, desc, in);
}

template <typename Float, typename Method>
infer_result<> infer_on(hip::queue& q, const descriptor<Float, Method>& desc,
                        const infer_input<>& in) {
    check_infer_preconditions(desc, in);
    if constexpr (std::is_same_v<Method, method::lloyd_csr>)
        return infer_csr_on<Float>(q, desc, in);
    else
        return infer_dense_on<Float>(q, desc, in);
}

}  // namespace detail
}  // namespace kmeans

// ---------------------------------------------------------------------------
// dal::train / dal::infer (train.hpp:26-59, infer.hpp)
// ---------------------------------------------------------------------------
template <typename Float, typename Method>
kmeans::train_result<> train(hip::queue& q, const kmeans::descriptor<Float, Method>& desc,
                             const kmeans::train_input<>& input) {
    return kmeans::detail::train_on(q, desc, input);
}
template <typename Float, typename Method>
kmeans::train_result<> train(hip::queue& q, const kmeans::descriptor<Float, Method>& desc,
                             const table& data, const table& initial_centroids = table{}) {
    return kmeans::detail::train_on(q, desc, kmeans::train_input<>(data, initial_centroids));
}
// Without a queue: the default device (0).  Preconditions are checked before
// the device is touched, as the reference checks them before dispatch.
template <typename Float, typename Method>
kmeans::train_result<> train(const kmeans::descriptor<Float, Method>& desc,
                             const kmeans::train_input<>& input) {
    kmeans::detail::check_train_preconditions(desc, input);
    return kmeans::detail::train_on(hip::default_queue(), desc, input);
}
template <typename Float, typename Method>
kmeans::train_result<> train(const kmeans::descriptor<Float, Method>& desc, const table& data,
                             const table& initial_centroids = table{}) {
    return train(desc, kmeans::train_input<>(data, initial_centroids));
}

template <typename Float, typename Method>
kmeans::infer_result<> infer(hip::queue& q, const kmeans::descriptor<Float, Method>& desc,
                             const kmeans::infer_input<>& input) {
    return kmeans::detail::infer_on(q, desc, input);
}
template <typename Float, typename Method>
kmeans::infer_result<> infer(hip::queue& q, const kmeans::descriptor<Float, Method>& desc,
                             const kmeans::model<>& m, const table& data) {
    return kmeans::detail::infer_on(q, desc, kmeans::infer_input<>(m, data));
}
template <typename Float, typename Method>
kmeans::infer_result<> infer(const kmeans::descriptor<Float, Method>& desc,
                             const kmeans::model<>& m, const table& data) {
    const kmeans::infer_input<> input(m, data);
    kmeans::detail::check_infer_preconditions(desc, input);
    return kmeans::detail::infer_on(hip::default_queue(), desc, input);
}

}  // namespace oneapi::dal
