// oneapi/dal/exceptions.hpp -- exception types of the drop-in C++ surface.
//
// Same names and hierarchy as the reference (cpp/oneapi/dal/exceptions.hpp:28-100):
// every type derives from oneapi::dal::exception and from the matching std::
// exception, so code that catches either keeps working.
#pragma once

#include <new>
#include <stdexcept>
#include <string>

namespace oneapi::dal {
inline namespace v1 {

class exception {
public:
    virtual ~exception() = default;
    virtual const char* what() const noexcept = 0;
};

class logic_error : public exception {};
class runtime_error : public exception {};

class invalid_argument : public logic_error, public std::invalid_argument {
public:
    using std::invalid_argument::invalid_argument;
    const char* what() const noexcept override { return std::invalid_argument::what(); }
};

class domain_error : public logic_error, public std::domain_error {
public:
    using std::domain_error::domain_error;
    const char* what() const noexcept override { return std::domain_error::what(); }
};

class out_of_range : public logic_error, public std::out_of_range {
public:
    using std::out_of_range::out_of_range;
    const char* what() const noexcept override { return std::out_of_range::what(); }
};

class unimplemented : public logic_error, public std::logic_error {
public:
    using std::logic_error::logic_error;
    const char* what() const noexcept override { return std::logic_error::what(); }
};

class unsupported_device : public logic_error, public std::logic_error {
public:
    using std::logic_error::logic_error;
    const char* what() const noexcept override { return std::logic_error::what(); }
};

class internal_error : public runtime_error, public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
    const char* what() const noexcept override { return std::runtime_error::what(); }
};

class bad_alloc : public exception, public std::bad_alloc {
public:
    const char* what() const noexcept override { return std::bad_alloc::what(); }
};

class host_bad_alloc : public bad_alloc {
public:
    const char* what() const noexcept override { return "Failed to allocate memory on host"; }
};

class device_bad_alloc : public bad_alloc {
public:
    const char* what() const noexcept override { return "Failed to allocate memory on device"; }
};

}  // namespace v1

// Messages used by the k-means path, worded as in the reference
// (cpp/oneapi/dal/detail/error_messages.cpp:65-66, 160-200).
namespace detail::error_messages {
inline const char* empty_set_of_result_options() { return "Empty set of result options"; }
inline const char* this_result_is_not_enabled_via_result_options() {
    return "This result is not enabled via result options";
}
inline const char* accuracy_threshold_lt_zero() { return "Accuracy_threshold is lower than zero"; }
inline const char* input_data_is_empty() { return "Input data is empty"; }
inline const char* iteration_count_lt_zero() { return "Iteration count is lower than zero"; }
inline const char* max_iteration_count_lt_zero() { return "Max iteration count lower than zero"; }
inline const char* cluster_count_leq_zero() { return "Cluster count is lower than or equal to zero"; }
inline const char* cluster_count_exceeds_data_row_count() {
    return "Cluster count exceeds data row count";
}
inline const char* cluster_count_gt_max_int32() {
    return "Cluster count is greater than max int32 value";
}
inline const char* row_count_gt_max_int32() { return "Row count is greater than max int32 value"; }
inline const char* input_initial_centroids_cc_neq_input_data_cc() {
    return "Input initial centroids column count is not equal to input data column count";
}
inline const char* input_initial_centroids_rc_neq_desc_cluster_count() {
    return "Input initial centroids row count is not equal to descriptor cluster count";
}
inline const char* input_model_centroids_are_empty() { return "Input model centroids are empty"; }
inline const char* input_model_centroids_cc_neq_input_data_cc() {
    return "Input model centroids column count is not equal to input data column count";
}
inline const char* input_model_centroids_rc_neq_desc_cluster_count() {
    return "Input model centroids row count is not equal to descriptor cluster count";
}
inline const char* objective_function_value_lt_zero() {
    return "Objective function value is lower than zero";
}
}  // namespace detail::error_messages

}  // namespace oneapi::dal
