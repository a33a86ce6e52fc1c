// km_args.hpp -- host-side precondition checks shared by the engine and the
// server.  Order and outcome follow cpp/oneapi/dal/algo/kmeans/detail/
// train_ops.hpp:40-61 and infer_ops.hpp:40-56; each failure maps onto the
// oneapi::dal exception type of the reference (exceptions.hpp:28-100).
#pragma once

#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>

#include "../../include/onedal_amd/kmeans_hip.h"

namespace km {

struct status_error : std::runtime_error {
    km_status status;
    status_error(km_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void check_train_args(const void* X, int64_t n, int64_t d, const void* C0, int64_t k,
                             int64_t max_iter, double thr, const void* C_out, const void* labels) {
    if (k <= 0) throw status_error(KM_DOMAIN_ERROR, "Cluster count is lower than or equal to zero");
    if (k > std::numeric_limits<int32_t>::max())
        throw status_error(KM_DOMAIN_ERROR, "Cluster count is greater than max int32 value");
    if (max_iter < 0) throw status_error(KM_DOMAIN_ERROR, "Max iteration count lower than zero");
    if (thr < 0.0) throw status_error(KM_DOMAIN_ERROR, "Accuracy_threshold is lower than zero");
    if (X == nullptr || n <= 0 || d <= 0) throw status_error(KM_DOMAIN_ERROR, "Input data is empty");
    if (n > std::numeric_limits<int32_t>::max())
        throw status_error(KM_DOMAIN_ERROR, "Row count is greater than max int32 value");
    if (k > n) throw status_error(KM_INVALID_ARGUMENT, "Cluster count exceeds data row count");
    if (C0 == nullptr)
        throw status_error(KM_UNIMPLEMENTED,
                           "Initial centroids are required: default k-means++ initialisation is "
                           "not part of this engine");
    if (C_out == nullptr || labels == nullptr)
        throw status_error(KM_INVALID_ARGUMENT, "Output buffers are required");
    if (d > 4096) throw status_error(KM_UNIMPLEMENTED, "column count above 4096 is not supported");
}

inline void check_infer_args(const void* X, int64_t n, int64_t d, const void* C, int64_t k,
                             uint32_t flags, const void* labels, const void* obj) {
    if (k <= 0) throw status_error(KM_DOMAIN_ERROR, "Cluster count is lower than or equal to zero");
    if (k > std::numeric_limits<int32_t>::max())
        throw status_error(KM_DOMAIN_ERROR, "Cluster count is greater than max int32 value");
    if (X == nullptr || n <= 0 || d <= 0) throw status_error(KM_DOMAIN_ERROR, "Input data is empty");
    if (C == nullptr) throw status_error(KM_DOMAIN_ERROR, "Input model centroids are empty");
    if (n > std::numeric_limits<int32_t>::max())
        throw status_error(KM_DOMAIN_ERROR, "Row count is greater than max int32 value");
    if (labels == nullptr) throw status_error(KM_INVALID_ARGUMENT, "Output buffers are required");
    if ((flags & KM_FLAG_EXACT_OBJECTIVE) && obj == nullptr)
        throw status_error(KM_INVALID_ARGUMENT, "Objective output is required");
    if (d > 4096) throw status_error(KM_UNIMPLEMENTED, "column count above 4096 is not supported");
}

// CSR arrays (csr_table: table/csr.hpp; the structure itself is checked on the
// device before use)
inline void check_csr_args(const void* values, const void* cols, const void* offs, int64_t nnz,
                           int32_t indexing) {
    if (indexing != KM_SPARSE_ZERO_BASED && indexing != KM_SPARSE_ONE_BASED)
        throw status_error(KM_INVALID_ARGUMENT, "Sparse indexing must be zero-based or one-based");
    if (nnz < 0) throw status_error(KM_INVALID_ARGUMENT, "Non-zero count is negative");
    if (offs == nullptr || (nnz > 0 && (values == nullptr || cols == nullptr)))
        throw status_error(KM_DOMAIN_ERROR, "Input data is empty");
}

}  // namespace km
