// oneapi/dal/table.hpp -- minimal homogeneous table of the drop-in C++ surface.
//
// Covers what the k-means path exchanges (reference cpp/oneapi/dal/table/
// homogen.hpp:48-160, row_accessor.hpp:70-135): row-major float / double /
// int32 data, shared on copy, held on the host or on a device (the reference's
// USM tables: homogen_table::wrap(queue, pointer, ...)), and
// row_accessor<const T>::pull with element-type conversion (pull(queue, ...)
// copies device rows to the host).
#pragma once

#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../onedal_amd/kmeans_hip.h"
#include "exceptions.hpp"

namespace oneapi::dal {

enum class data_type { int32, float32, float64 };

namespace detail {
template <typename T>
constexpr data_type data_type_of() {
    if constexpr (std::is_same_v<T, float>) return data_type::float32;
    else if constexpr (std::is_same_v<T, double>) return data_type::float64;
    else {
        static_assert(std::is_same_v<T, std::int32_t>, "tables hold float, double or int32");
        return data_type::int32;
    }
}
inline std::size_t size_of(data_type t) { return t == data_type::float64 ? 8 : 4; }
}  // namespace detail

// Host array with shared ownership (the reference's dal::array, reduced).
template <typename T>
class array {
public:
    array() = default;
    explicit array(std::int64_t count) : data_(std::make_shared<std::vector<T>>(count)) {}
    static array zeros(std::int64_t count) { return array(count); }
    const T* get_data() const { return data_ ? data_->data() : nullptr; }
    T* get_mutable_data() { return data_ ? data_->data() : nullptr; }
    std::int64_t get_count() const { return data_ ? std::int64_t(data_->size()) : 0; }
    const T& operator[](std::int64_t i) const { return (*data_)[i]; }

private:
    std::shared_ptr<std::vector<T>> data_;
};

// table/common.hpp:37
enum class sparse_indexing { zero_based, one_based };

class table {
public:
    table() = default;
    bool has_data() const {
        return rows_ > 0 && cols_ > 0 && (kind_ == 1 ? csr_offs_ != nullptr : data_ != nullptr);
    }
    // 0: homogen_table, 1: csr_table (homogen_table::kind(), csr_table::kind())
    std::int64_t get_kind() const { return kind_; }
    std::int64_t get_row_count() const { return rows_; }
    std::int64_t get_column_count() const { return cols_; }
    data_type get_data_type() const { return dtype_; }
    const void* get_raw_data() const { return data_.get(); }
    // device tables: the HIP device holding the data (-1: host memory)
    bool is_device() const { return device_ >= 0; }
    int get_device() const { return device_; }

protected:
    std::shared_ptr<const void> data_;
    std::int64_t rows_ = 0, cols_ = 0;
    data_type dtype_ = data_type::float32;
    int device_ = -1;
    // CSR tables: column indices and row offsets (int64), stored entry count and
    // indexing; kept in the base so a table passed by value keeps them
    std::int64_t kind_ = 0;
    std::shared_ptr<const void> csr_cols_, csr_offs_;
    std::int64_t nnz_ = 0;
    sparse_indexing indexing_ = sparse_indexing::one_based;
};

class homogen_table : public table {
public:
    homogen_table() = default;
    static std::int64_t kind() { return 0; }

    // Non-owning view of row-major host data (the caller keeps it alive).
    template <typename T>
    static homogen_table wrap(const T* data, std::int64_t rows, std::int64_t cols) {
        homogen_table t;
        t.data_ = std::shared_ptr<const void>(data, [](const void*) {});
        t.rows_ = rows;
        t.cols_ = cols;
        t.dtype_ = detail::data_type_of<T>();
        return t;
    }

    // Non-owning view of row-major data in device memory of the queue's device
    // (the reference's homogen_table::wrap(queue, usm_pointer, ...)).
    template <typename Queue, typename T>
    static homogen_table wrap(const Queue& q, const T* device_data, std::int64_t rows, std::int64_t cols) {
        homogen_table t = wrap(device_data, rows, cols);
        t.device_ = q.get_device();
        return t;
    }

    // Device data whose lifetime the table shares (results of device calls).
    template <typename T>
    static homogen_table wrap_device(std::shared_ptr<const T> data, std::int64_t rows, std::int64_t cols,
                                     int device) {
        homogen_table t;
        t.data_ = std::shared_ptr<const void>(data, data.get());
        t.rows_ = rows;
        t.cols_ = cols;
        t.dtype_ = detail::data_type_of<T>();
        t.device_ = device;
        return t;
    }

    // Shares ownership of a dal::array.
    template <typename T>
    static homogen_table wrap(const array<T>& a, std::int64_t rows, std::int64_t cols) {
        homogen_table t;
        auto holder = std::make_shared<array<T>>(a);
        t.data_ = std::shared_ptr<const void>(holder, holder->get_data());
        t.rows_ = rows;
        t.cols_ = cols;
        t.dtype_ = detail::data_type_of<T>();
        return t;
    }
};

// Compressed sparse rows (reference table/csr.hpp:30-352): values, column
// indices and row offsets, zero- or one-based (one-based by default, as there).
// wrap() makes non-owning views of host arrays, or of device arrays on the
// queue's device.
class csr_table : public table {
public:
    static std::int64_t kind() { return 1; }
    csr_table() { kind_ = 1; }
    // a table of CSR kind viewed as a csr_table (the reference's static_cast)
    explicit csr_table(const table& t) : table(t) {
        if (t.get_kind() != kind()) throw invalid_argument("The table is not a CSR table");
    }

    template <typename T>
    static csr_table wrap(const T* data, const std::int64_t* column_indices,
                          const std::int64_t* row_offsets, std::int64_t row_count,
                          std::int64_t column_count,
                          sparse_indexing indexing = sparse_indexing::one_based) {
        csr_table t = make<T>(data, column_indices, row_offsets, row_count, column_count, indexing);
        t.nnz_ = row_count > 0 && row_offsets ? row_offsets[row_count] - row_offsets[0] : 0;
        return t;
    }

    template <typename Queue, typename T>
    static csr_table wrap(const Queue& q, const T* data, const std::int64_t* column_indices,
                          const std::int64_t* row_offsets, std::int64_t row_count,
                          std::int64_t column_count,
                          sparse_indexing indexing = sparse_indexing::one_based) {
        csr_table t = make<T>(data, column_indices, row_offsets, row_count, column_count, indexing);
        t.device_ = q.get_device();
        std::int64_t ends[2] = {0, 0};
        if (row_count > 0 && row_offsets) {
            for (int e = 0; e < 2; ++e) {
                const km_status st = km_memcpy(q.get_context(), &ends[e], row_offsets + (e ? row_count : 0),
                                               sizeof(std::int64_t), KM_COPY_DEVICE_TO_HOST);
                if (st != KM_OK)
                    throw internal_error(std::string("device copy failed: ") + km_last_error(q.get_context()));
            }
        }
        t.nnz_ = ends[1] - ends[0];
        return t;
    }

    std::int64_t get_non_zero_count() const { return nnz_; }
    sparse_indexing get_indexing() const { return indexing_; }
    const void* get_data() const { return data_.get(); }
    const std::int64_t* get_column_indices() const {
        return static_cast<const std::int64_t*>(csr_cols_.get());
    }
    const std::int64_t* get_row_offsets() const {
        return static_cast<const std::int64_t*>(csr_offs_.get());
    }

private:
    template <typename T>
    static csr_table make(const T* data, const std::int64_t* cols, const std::int64_t* offs,
                          std::int64_t rows, std::int64_t ncols, sparse_indexing indexing) {
        csr_table t;
        t.data_ = std::shared_ptr<const void>(data, [](const void*) {});
        t.csr_cols_ = std::shared_ptr<const void>(cols, [](const void*) {});
        t.csr_offs_ = std::shared_ptr<const void>(offs, [](const void*) {});
        t.rows_ = rows;
        t.cols_ = ncols;
        t.dtype_ = detail::data_type_of<T>();
        t.indexing_ = indexing;
        return t;
    }
};

struct range {
    std::int64_t start_idx = 0;
    std::int64_t end_idx = -1;  // -1: up to the last row
};

template <typename T>
class row_accessor {
    static_assert(std::is_const_v<T>, "row_accessor gives read-only access");
    using V = std::remove_const_t<T>;

public:
    explicit row_accessor(const table& t) : t_(t) {
        if (t.get_kind() != 0) throw invalid_argument("row_accessor reads homogeneous tables");
    }

    // Rows [r.start_idx, r.end_idx) of a host table as a host array of V,
    // converted if needed.  Device tables need the queue overload.
    array<V> pull(const range& r = {}) const {
        if (t_.is_device())
            throw invalid_argument("The table holds device data: pull it with a queue");
        std::int64_t start, count;
        const char* base = rows_of(r, start, count);
        return convert_all(base, count);
    }

    // The same rows copied from a device table through the queue's context
    // (host tables are read directly).
    template <typename Queue>
    array<V> pull(const Queue& q, const range& r = {}) const {
        if (!t_.is_device()) return pull(r);
        if (q.get_device() != t_.get_device())
            throw invalid_argument("The queue's device does not hold the table");
        std::int64_t start, count;
        const char* base = rows_of(r, start, count);
        std::vector<char> host(std::size_t(count) * detail::size_of(t_.get_data_type()));
        const km_status st = km_memcpy(q.get_context(), host.data(), base, std::int64_t(host.size()),
                                       KM_COPY_DEVICE_TO_HOST);
        if (st != KM_OK) throw internal_error(std::string("device copy failed: ") + km_last_error(q.get_context()));
        return convert_all(host.data(), count);
    }

private:
    // first row and element count of r; pointer to its first element
    const char* rows_of(const range& r, std::int64_t& start, std::int64_t& count) const {
        const std::int64_t end = r.end_idx < 0 ? t_.get_row_count() : r.end_idx;
        if (r.start_idx < 0 || end > t_.get_row_count() || r.start_idx > end)
            throw out_of_range("Row range is out of table bounds");
        const std::int64_t cols = t_.get_column_count();
        start = r.start_idx;
        count = (end - r.start_idx) * cols;
        return static_cast<const char*>(t_.get_raw_data()) + r.start_idx * cols * detail::size_of(t_.get_data_type());
    }

    array<V> convert_all(const char* base, std::int64_t count) const {
        array<V> out(count);
        switch (t_.get_data_type()) {
            case data_type::float32: convert(reinterpret_cast<const float*>(base), out, count); break;
            case data_type::float64: convert(reinterpret_cast<const double*>(base), out, count); break;
            case data_type::int32: convert(reinterpret_cast<const std::int32_t*>(base), out, count); break;
        }
        return out;
    }

    template <typename S>
    static void convert(const S* src, array<V>& out, std::int64_t count) {
        V* dst = out.get_mutable_data();
        if constexpr (std::is_same_v<S, V>) {
            if (count) std::memcpy(dst, src, sizeof(V) * count);
        } else {
            for (std::int64_t i = 0; i < count; ++i) dst[i] = V(src[i]);
        }
    }
    table t_;
};

}  // namespace oneapi::dal
