// oneapi/dal/archive.hpp -- binary archives for model serialization.
//
// The reference serializes models through dal::detail::serialize /
// deserialize over output/input archives (cpp/oneapi/dal/archive.hpp,
// backend/serialization.hpp), tagging each model implementation with a
// serialization id (kmeans: kmeans_clustering_model_impl_id = 8010000000,
// backend/serialization.hpp:194).  This surface keeps the calls and the id; the
// byte layout is this library's own (a stream of fixed-width fields), not the
// reference's, so archives are exchanged between builds of this surface only.
#pragma once

#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "exceptions.hpp"

namespace oneapi::dal::detail {

using byte_t = std::uint8_t;

class binary_output_archive {
public:
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

    template <typename T>
    void operator()(T& value) {
        static_assert(std::is_trivially_copyable_v<T>, "archives hold trivially copyable fields");
        read(&value, sizeof(T));
    }
    void read(void* dst, std::int64_t bytes) {
        if (bytes < 0 || pos_ + bytes > size_) throw invalid_argument("Archive is truncated");
        if (bytes) std::memcpy(dst, data_ + pos_, std::size_t(bytes));
        pos_ += bytes;
    }

private:
    const byte_t* data_;
    std::int64_t size_;
    std::int64_t pos_ = 0;
};

// dal::detail::serialize / deserialize (cpp/oneapi/dal/archive.hpp)
template <typename Object>
void serialize(const Object& object, binary_output_archive& archive) {
    object.serialize(archive);
}

template <typename Object>
void deserialize(Object& object, binary_input_archive& archive) {
    object.deserialize(archive);
}

}  // namespace oneapi::dal::detail
