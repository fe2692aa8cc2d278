// bld_fact.hpp — W, H, L from S (role of src/bld/bld_fact.hpp:36-80): the
// axis sums come from the GPU, the O((x+y)z) normalisation runs here.
#pragma once

#include <tuple>

#include "../alloc_model/alloc_model_params.hpp"
#include "../util/util.hpp"

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
    const auto sums = util::tensor_sums(S);
    const tensor_t<T, 2>& S_pjk = sums[0];
    const tensor_t<T, 2>& S_ipk = sums[1];
    matrix_t<T> W(x, z), H(z, y);
    vector_t<T> L(1, y);
    for (long i = 0; i < x; ++i)
        for (long k = 0; k < z; ++k) W(i, k) = p.alpha[i] + S_ipk(i, k) - 1;
    for (long k = 0; k < z; ++k)
        for (long j = 0; j < y; ++j) H(k, j) = p.beta[k] + S_pjk(j, k) - 1;
    for (long j = 0; j < y; ++j) {
        T s = 0;
        for (long k = 0; k < z; ++k) s += S_pjk(j, k);
        L(0, j) = (p.a + s - 1) / (p.b + 1 + eps);
    }
    for (long k = 0; k < z; ++k) {
        T cs = 0;
        for (long i = 0; i < x; ++i) cs += W(i, k);
        cs += eps;
        for (long i = 0; i < x; ++i) W(i, k) /= cs;
    }
    for (long j = 0; j < y; ++j) {
        T cs = 0;
        for (long k = 0; k < z; ++k) cs += H(k, j);
        cs += eps;
        for (long k = 0; k < z; ++k) H(k, j) /= cs;
    }
    return std::make_tuple(W, H, L);
}

}  // namespace bld
}  // namespace bnmf_algs
