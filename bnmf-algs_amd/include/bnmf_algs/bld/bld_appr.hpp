// bld_appr.hpp — drop-in for bld::bld_appr (reference
// src/bld/bld_appr.hpp:286-290): the approximate BLD solver that keeps S in
// factored form S_ijk = X_ij nu_ik mu_kj / (nu mu + eps)_ij and updates nu
// and mu multiplicatively, run on the GPU through the C ABI
// (include/bnmf_hip.h). Returns (S, nu, mu) as the reference does. The
// initial factors are drawn from Dirichlet(1_z) (rows of nu, columns of mu)
// from a counter-based stream; the overload taking nu0 / mu0 starts there.
#pragma once

#include <cstdint>
#include <tuple>

#include "../alloc_model/alloc_model_params.hpp"
#include "../defs.hpp"
#include "../detail.hpp"
#include "bld_mult.hpp"

namespace bnmf_algs {
namespace bld {

namespace details {
template <typename T, typename Scalar>
std::tuple<tensor_t<T, 3>, matrix_t<T>, matrix_t<T>> solve_appr(const matrix_t<T>& X, size_t z,
                                                                const alloc_model::Params<Scalar>& p,
                                                                const T* nu0, const T* mu0,
                                                                size_t max_iter, double eps,
                                                                uint64_t seed) {
    check_bld_params(X, z, p);
    const size_t x = static_cast<size_t>(X.rows()), y = static_cast<size_t>(X.cols());
    tensor_t<T, 3> S(static_cast<long>(x), static_cast<long>(y), static_cast<long>(z));
    matrix_t<T> nu(static_cast<long>(x), static_cast<long>(z));
    matrix_t<T> mu(static_cast<long>(z), static_cast<long>(y));
    const std::vector<double> alpha = bnmf_algs::details::to_double(p.alpha);
    const std::vector<double> beta = bnmf_algs::details::to_double(p.beta);
    bnmf_status st;
    if constexpr (bnmf_algs::details::dtype_of<T>() == BNMF_F64)
        st = bnmf_bld_appr_f64(X.data(), x, y, z, alpha.data(), beta.data(), max_iter, eps, nu0, mu0,
                               seed, S.data(), nu.data(), mu.data());
    else
        st = bnmf_bld_appr_f32(X.data(), x, y, z, alpha.data(), beta.data(), max_iter, eps, nu0, mu0,
                               seed, S.data(), nu.data(), mu.data());
    bnmf_algs::details::check_status(st, "bld_appr");
    return std::make_tuple(S, nu, mu);
}
}  // namespace details

template <typename T, typename Scalar>
std::tuple<tensor_t<T, 3>, matrix_t<T>, matrix_t<T>> bld_appr(const matrix_t<T>& X, size_t z,
                                                              const alloc_model::Params<Scalar>& model_params,
                                                              size_t max_iter = 1000, double eps = 1e-50) {
    return details::solve_appr<T, Scalar>(X, z, model_params, nullptr, nullptr, max_iter, eps,
                                          kDefaultInitSeed);
}

// Starts from the caller's nu0 (x*z, rows summing to 1) and mu0 (z*y,
// columns summing to 1).
template <typename T, typename Scalar>
std::tuple<tensor_t<T, 3>, matrix_t<T>, matrix_t<T>> bld_appr(const matrix_t<T>& X, size_t z,
                                                              const alloc_model::Params<Scalar>& model_params,
                                                              const matrix_t<T>& nu0, const matrix_t<T>& mu0,
                                                              size_t max_iter = 1000, double eps = 1e-50) {
    BNMF_ASSERT(nu0.rows() == X.rows() && static_cast<size_t>(nu0.cols()) == z,
                "nu0 must have shape (x, z)");
    BNMF_ASSERT(static_cast<size_t>(mu0.rows()) == z && mu0.cols() == X.cols(),
                "mu0 must have shape (z, y)");
    return details::solve_appr<T, Scalar>(X, z, model_params, nu0.data(), mu0.data(), max_iter, eps,
                                          kDefaultInitSeed);
}

}  // namespace bld
}  // namespace bnmf_algs
