// oneapi/dal/archive.hpp -- binary archives for model serialization, in the
// reference's byte format.
//
// Format (reference cpp/oneapi/dal/detail/archives.hpp:23-128 and
// detail/serialization.hpp:379-448): dal::detail::serialize writes the archive
// prologue, a uint32 magic 0x4441414F, then the object; every field is its
// native (little-endian) bytes with no type tags.  bool fields are one byte
// (trivial_serialization_type<bool> = uint8, serialization.hpp:88-91) and enums
// their underlying integer.  A polymorphic member is a one-byte "present" flag,
// then its uint64 serialization id (backend/serialization.hpp:155-194), then its
// fields (serialize_polymorphic, serialization.hpp:379-389).  An array is its
// array_id, the element data_type code (int32), the size in bytes (int64) and the
// bytes (detail/array_impl.cpp:26-37).  A homogeneous table is (metadata, data
// bytes as a uint8 array, row count, column count, layout)
// (table/backend/homogen_table_impl.hpp:136-138), its metadata the per-column
// data types and feature types as two int32 arrays (table/common.cpp:133-141).
#pragma once

#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "exceptions.hpp"

namespace oneapi::dal::detail {

using byte_t = std::uint8_t;

constexpr std::uint32_t binary_archive_magic = 0x4441414F;

inline const char* archive_content_does_not_match_type() {
    return "Archive content does not match requested type";
}

// serialization ids (backend/serialization.hpp:155-194)
namespace serialization_ids {
constexpr std::uint64_t array_id = 1000000000;
constexpr std::uint64_t empty_table_metadata_id = 2000000000;
constexpr std::uint64_t simple_table_metadata_id = 2010000000;
constexpr std::uint64_t empty_table_id = 2020000000;
constexpr std::uint64_t homogen_table_id = 2030000000;
constexpr std::uint64_t kmeans_clustering_model_impl_id = 8010000000;
}  // namespace serialization_ids

// data_type codes as archived (enum class data_type : int32, common.hpp:97-109)
namespace archived_type {
constexpr std::int32_t int32 = 0x0004;
constexpr std::int32_t uint8 = 0x0010;
constexpr std::int32_t float32 = 0x0100;
constexpr std::int32_t float64 = 0x0200;
}  // namespace archived_type

// feature_type / data_layout values (table/common.hpp:35-36)
constexpr std::int32_t feature_type_ordinal = 1;
constexpr std::int32_t feature_type_ratio = 3;
constexpr std::int32_t data_layout_row_major = 1;

class binary_output_archive {
public:
    void prologue() { (*this)(binary_archive_magic); }
    void epilogue() {}
    // a trivially copyable value, as its bytes
    template <typename T>
    void operator()(const T& value) {
        static_assert(std::is_trivially_copyable_v<T>, "archives hold trivially copyable fields");
        write(&value, sizeof(T));
    }
    void write(const void* data, std::int64_t bytes) {
        const byte_t* p = static_cast<const byte_t*>(data);
        buf_.insert(buf_.end(), p, p + bytes);
    }
    std::int64_t get_size() const { return std::int64_t(buf_.size()); }
    void to_array(byte_t* dst, std::int64_t size) const {
        if (size < get_size()) throw invalid_argument("Destination is smaller than the archive");
        if (!buf_.empty()) std::memcpy(dst, buf_.data(), buf_.size());
    }
    std::vector<byte_t> to_array() const { return buf_; }

private:
    std::vector<byte_t> buf_;
};

class binary_input_archive {
public:
    binary_input_archive(const byte_t* data, std::int64_t size) : data_(data), size_(size) {}

    void prologue() {
        std::uint32_t magic = 0;
        (*this)(magic);
        if (magic != binary_archive_magic) throw invalid_argument(archive_content_does_not_match_type());
    }
    void epilogue() {}
    template <typename T>
    void operator()(T& value) {
        static_assert(std::is_trivially_copyable_v<T>, "archives hold trivially copyable fields");
        read(&value, sizeof(T));
    }
    template <typename T>
    T pop() {
        T v{};
        (*this)(v);
        return v;
    }
    void read(void* dst, std::int64_t bytes) {
        if (bytes < 0 || bytes > remaining()) throw invalid_argument(archive_content_does_not_match_type());
        if (bytes) std::memcpy(dst, data_ + pos_, std::size_t(bytes));
        pos_ += bytes;
    }
    std::int64_t remaining() const { return size_ - pos_; }

private:
    const byte_t* data_;
    std::int64_t size_;
    std::int64_t pos_ = 0;
};

// One array field (detail/array_impl.cpp:26-37).
inline void write_array(binary_output_archive& ar, std::int32_t dtype_code, const void* bytes,
                        std::int64_t size_in_bytes) {
    ar(serialization_ids::array_id);
    ar(dtype_code);
    ar(size_in_bytes);
    if (size_in_bytes > 0) ar.write(bytes, size_in_bytes);
}

// Reads an array field of element code `dtype_code` and element size `elem`
// (detail/array_impl.cpp:39-67); the byte count is checked against what the
// archive still holds before anything is allocated.
inline std::vector<byte_t> read_array(binary_input_archive& ar, std::int32_t dtype_code,
                                      std::int64_t elem) {
    if (ar.pop<std::uint64_t>() != serialization_ids::array_id ||
        ar.pop<std::int32_t>() != dtype_code)
        throw invalid_argument(archive_content_does_not_match_type());
    const std::int64_t size = ar.pop<std::int64_t>();
    if (size < 0 || size % elem != 0 || size > ar.remaining())
        throw invalid_argument(archive_content_does_not_match_type());
    std::vector<byte_t> out(static_cast<std::size_t>(size));
    ar.read(out.data(), size);
    return out;
}

// A table member: an empty table (no payload), or a row-major homogeneous table
// of `rows` x `cols` elements of archived type `dtype_code` held in host memory.
inline void write_table(binary_output_archive& ar, std::int64_t rows, std::int64_t cols,
                        std::int32_t dtype_code, std::int64_t elem, const void* host_bytes) {
    ar(std::uint8_t(1));
    if (rows <= 0 || cols <= 0) {
        ar(serialization_ids::empty_table_id);
        return;
    }
    ar(serialization_ids::homogen_table_id);
    // metadata: simple_metadata_impl (dtypes, ftypes)
    ar(std::uint8_t(1));
    ar(serialization_ids::simple_table_metadata_id);
    const std::int32_t ftype = dtype_code == archived_type::float32 || dtype_code == archived_type::float64
                                   ? feature_type_ratio
                                   : feature_type_ordinal;
    std::vector<std::int32_t> dtypes(static_cast<std::size_t>(cols), dtype_code);
    std::vector<std::int32_t> ftypes(static_cast<std::size_t>(cols), ftype);
    write_array(ar, archived_type::int32, dtypes.data(), cols * 4);
    write_array(ar, archived_type::int32, ftypes.data(), cols * 4);
    // data bytes, row count, column count, layout
    write_array(ar, archived_type::uint8, host_bytes, rows * cols * elem);
    ar(rows);
    ar(cols);
    ar(data_layout_row_major);
}

struct archived_table {
    std::int64_t rows = 0, cols = 0;
    std::int32_t dtype_code = 0;  // 0: empty table
    std::vector<byte_t> bytes;
};

inline archived_table read_table(binary_input_archive& ar) {
    archived_table t;
    if (ar.pop<std::uint8_t>() != 1) throw invalid_argument(archive_content_does_not_match_type());
    const std::uint64_t id = ar.pop<std::uint64_t>();
    if (id == serialization_ids::empty_table_id) return t;
    if (id != serialization_ids::homogen_table_id) throw invalid_argument(archive_content_does_not_match_type());
    if (ar.pop<std::uint8_t>() != 1) throw invalid_argument(archive_content_does_not_match_type());
    const std::uint64_t mid = ar.pop<std::uint64_t>();
    std::vector<byte_t> dtypes, ftypes;
    if (mid == serialization_ids::simple_table_metadata_id) {
        dtypes = read_array(ar, archived_type::int32, 4);
        ftypes = read_array(ar, archived_type::int32, 4);
    } else if (mid != serialization_ids::empty_table_metadata_id) {
        throw invalid_argument(archive_content_does_not_match_type());
    }
    t.bytes = read_array(ar, archived_type::uint8, 1);
    t.rows = ar.pop<std::int64_t>();
    t.cols = ar.pop<std::int64_t>();
    const std::int32_t layout = ar.pop<std::int32_t>();
    if (t.rows == 0 && t.cols == 0 && t.bytes.empty()) return t;  // an empty homogen table
    if (t.rows <= 0 || t.cols <= 0 || layout != data_layout_row_major || dtypes.size() < 4 ||
        std::int64_t(dtypes.size()) != 4 * t.cols || ftypes.size() != dtypes.size())
        throw invalid_argument(archive_content_does_not_match_type());
    std::memcpy(&t.dtype_code, dtypes.data(), 4);
    for (std::size_t i = 4; i < dtypes.size(); i += 4)
        if (std::memcmp(dtypes.data() + i, dtypes.data(), 4) != 0)
            throw invalid_argument(archive_content_does_not_match_type());
    std::int64_t elem = 0;
    if (t.dtype_code == archived_type::float32 || t.dtype_code == archived_type::int32) elem = 4;
    if (t.dtype_code == archived_type::float64) elem = 8;
    if (elem == 0 || t.rows > std::int64_t(t.bytes.size()) / elem / t.cols ||
        t.rows * t.cols * elem != std::int64_t(t.bytes.size()))
        throw invalid_argument(archive_content_does_not_match_type());
    return t;
}

// dal::detail::serialize / deserialize (serialization.hpp:433-448)
template <typename Object>
void serialize(const Object& object, binary_output_archive& archive) {
    archive.prologue();
    object.serialize(archive);
    archive.epilogue();
}

template <typename Object>
void deserialize(Object& object, binary_input_archive& archive) {
    archive.prologue();
    object.deserialize(archive);
    archive.epilogue();
}

}  // namespace oneapi::dal::detail
