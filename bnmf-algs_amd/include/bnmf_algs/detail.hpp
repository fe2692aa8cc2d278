// detail.hpp — glue between the drop-in C++ templates and the C ABI
// (include/bnmf_hip.h): type dispatch and status-to-exception mapping.
#pragma once

#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "bnmf_hip.h"

namespace bnmf_algs {
namespace details {

// Parameter errors become std::invalid_argument (the reference's
// BNMF_ASSERT under BUILD_TESTING, src/defs.hpp:12-16); device and memory
// failures become std::runtime_error.
inline void check_status(bnmf_status st, const char* what) {
    if (st == BNMF_OK) return;
    const std::string msg = std::string(what) + ": " + bnmf_last_error();
    if (st == BNMF_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

template <typename T>
constexpr int dtype_of() {
    static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                  "the GPU solver computes in float or double");
    return std::is_same<T, float>::value ? BNMF_F32 : BNMF_F64;
}

template <typename Scalar>
std::vector<double> to_double(const std::vector<Scalar>& v) {
    return std::vector<double>(v.begin(), v.end());
}

}  // namespace details
}  // namespace bnmf_algs
