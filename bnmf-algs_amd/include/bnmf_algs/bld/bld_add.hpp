// bld_add.hpp — drop-in for bld::bld_add (reference src/bld/bld_add.hpp:69-72):
// additive gradient ascent on Z = log S with S = X * softmax_k(Z), run by the
// fused HIP solver through the C ABI (include/bnmf_hip.h). T = double runs
// every step in double, T = float in float; every digamma is psi_appr, as in
// the reference. The initial tensor is drawn as bld_mult's (S_ij: = X_ij *
// Dirichlet(1_z), the reference's inline draw at :76-91) from a
// counter-based stream; the overload that takes S0 starts from the caller's.
#pragma once

#include <cstdint>

#include "../alloc_model/alloc_model_params.hpp"
#include "../defs.hpp"
#include "../detail.hpp"
#include "bld_mult.hpp"

namespace bnmf_algs {
namespace bld {

namespace details {
template <typename T, typename Scalar>
tensor_t<T, 3> solve_add(const matrix_t<T>& X, size_t z, const alloc_model::Params<Scalar>& p,
                         const T* S0, size_t max_iter, double eps, uint64_t seed) {
    check_bld_params(X, z, p);
    const size_t x = static_cast<size_t>(X.rows()), y = static_cast<size_t>(X.cols());
    tensor_t<T, 3> S(static_cast<long>(x), static_cast<long>(y), static_cast<long>(z));
    const std::vector<double> alpha = bnmf_algs::details::to_double(p.alpha);
    const std::vector<double> beta = bnmf_algs::details::to_double(p.beta);
    bnmf_status st;
    if constexpr (bnmf_algs::details::dtype_of<T>() == BNMF_F64)
        st = bnmf_bld_add_f64(X.data(), x, y, z, alpha.data(), beta.data(), max_iter, eps, S0,
                              seed, S.data());
    else
        st = bnmf_bld_add_f32(X.data(), x, y, z, alpha.data(), beta.data(), max_iter, eps, S0,
                              seed, S.data());
    bnmf_algs::details::check_status(st, "bld_add");
    return S;
}
}  // namespace details

template <typename T, typename Scalar>
tensor_t<T, 3> bld_add(const matrix_t<T>& X, size_t z,
                       const alloc_model::Params<Scalar>& model_params, size_t max_iter = 1000,
                       double eps = 1e-50) {
    return details::solve_add<T, Scalar>(X, z, model_params, nullptr, max_iter, eps,
                                         kDefaultInitSeed);
}

// Starts from the caller's S0 (x*y*z) instead of a drawn one.
template <typename T, typename Scalar>
tensor_t<T, 3> bld_add(const matrix_t<T>& X, size_t z,
                       const alloc_model::Params<Scalar>& model_params, const tensor_t<T, 3>& S0,
                       size_t max_iter = 1000, double eps = 1e-50) {
    BNMF_ASSERT(S0.dimension(0) == X.rows() && S0.dimension(1) == X.cols() &&
                    static_cast<size_t>(S0.dimension(2)) == z,
                "S0 must have shape (x, y, z)");
    return details::solve_add<T, Scalar>(X, z, model_params, S0.data(), max_iter, eps,
                                         kDefaultInitSeed);
}

}  // namespace bld
}  // namespace bnmf_algs
