// bld_fact.hpp — W, H, L from S (src/bld/bld_fact.hpp:36-80): the axis sums
// and the normalisation run on the GPU (bnmf_bld_fact); results in T.
#pragma once

#include <tuple>
#include <vector>

#include "../alloc_model/alloc_model_params.hpp"
#include "../defs.hpp"
#include "../detail.hpp"

namespace bnmf_algs {
namespace bld {

template <typename T, typename Scalar>
std::tuple<matrix_t<T>, matrix_t<T>, vector_t<T>> bld_fact(const tensor_t<T, 3>& S,
                                                           const alloc_model::Params<Scalar>& p,
                                                           double eps = 1e-50) {
    const long x = S.dimension(0), y = S.dimension(1), z = S.dimension(2);
    BNMF_ASSERT(p.alpha.size() == static_cast<size_t>(x),
                "Number of alpha parameters must be equal to S.dimension(0)");
    BNMF_ASSERT(p.beta.size() == static_cast<size_t>(z), "Number of beta parameters must be equal to z");
    const std::vector<double> alpha = bnmf_algs::details::to_double(p.alpha);
    const std::vector<double> beta = bnmf_algs::details::to_double(p.beta);
    std::vector<double> w((size_t)(x * z)), h((size_t)(z * y)), l((size_t)y);
    bnmf_algs::details::check_status(
        bnmf_bld_fact(bnmf_algs::details::dtype_of<T>(), S.data(), (size_t)x, (size_t)y, (size_t)z,
                      alpha.data(), beta.data(), (double)p.a, (double)p.b, eps, w.data(), h.data(),
                      l.data()),
        "bld_fact");
    matrix_t<T> W(x, z), H(z, y);
    vector_t<T> L(1, y);
    for (long t = 0; t < x * z; ++t) W.data()[t] = static_cast<T>(w[(size_t)t]);
    for (long t = 0; t < z * y; ++t) H.data()[t] = static_cast<T>(h[(size_t)t]);
    for (long t = 0; t < y; ++t) L.data()[t] = static_cast<T>(l[(size_t)t]);
    return std::make_tuple(W, H, L);
}

}  // namespace bld
}  // namespace bnmf_algs
