// util.hpp — helpers of the drop-in API: tensor axis sums on the GPU (role
// of cuda::tensor_sums, src/cuda/tensor_ops.hpp:40-42) and sparseness
// (util::sparseness, src/util/util.hpp:270-295).
#pragma once

#include <array>
#include <cmath>
#include <limits>

#include "../defs.hpp"
#include "../detail.hpp"

namespace bnmf_algs {
namespace util {

// {S_pjk (y*z), S_ipk (x*z), S_ijp (x*y)}: sums over axes 0, 1 and 2.
template <typename T>
std::array<tensor_t<T, 2>, 3> tensor_sums(const tensor_t<T, 3>& S) {
    const long x = S.dimension(0), y = S.dimension(1), z = S.dimension(2);
    std::array<tensor_t<T, 2>, 3> out{tensor_t<T, 2>(y, z), tensor_t<T, 2>(x, z), tensor_t<T, 2>(x, y)};
    bnmf_algs::details::check_status(
        bnmf_tensor_sums(bnmf_algs::details::dtype_of<T>(), S.data(), (size_t)x, (size_t)y, (size_t)z,
                         out[0].data(), out[1].data(), out[2].data()),
        "tensor_sums");
    return out;
}

// (sqrt(xyz) - S_+++ / ||S||_F) / (sqrt(xyz) - 1); DBL_MAX for a zero tensor.
template <typename T>
double sparseness(const tensor_t<T, 3>& S) {
    const long x = S.dimension(0), y = S.dimension(1), z = S.dimension(2);
    double sum = 0, sq = 0;
    const T* d = S.data();
    for (long t = 0; t < x * y * z; ++t) {
        sum += d[t];
        sq += (double)d[t] * d[t];
    }
    if (sq < std::numeric_limits<double>::epsilon()) return std::numeric_limits<double>::max();
    const double axis_mult = std::sqrt((double)(x * y * z));
    return (axis_mult - sum / std::sqrt(sq)) / (axis_mult - 1);
}

}  // namespace util
}  // namespace bnmf_algs
