// bld_mult.hpp — drop-in for bld::bld_mult and bld::bld_mult_cuda
// (reference src/bld/bld_mult/bld_mult.hpp:82-86 and :225-229). Both run the
// fused HIP solver through the C ABI (include/bnmf_hip.h):
//   T = double: the plain FP64 form (exact digamma unless use_psi_appr);
//   T = float : the FP32-bulk form.
// The initial tensor is drawn as init_S does (S_ij: = X_ij * Dirichlet(1_z),
// bld_mult_funcs.hpp:36-74) from a counter-based stream; the overloads that
// take S0 start from the caller's tensor instead.
#pragma once

#include <cstdint>

#include "../alloc_model/alloc_model_params.hpp"
#include "../defs.hpp"
#include "../detail.hpp"

namespace bnmf_algs {
namespace bld {

// Seed of the S0 stream used when no S0 is given.
constexpr uint64_t kDefaultInitSeed = 0;

namespace details {
// check_bld_params (src/bld/util_details.hpp:20-28); X >= 0 is checked by
// the C ABI.
template <typename T, typename Scalar>
void check_bld_params(const matrix_t<T>& X, size_t z, const alloc_model::Params<Scalar>& p) {
    BNMF_ASSERT(p.alpha.size() == static_cast<size_t>(X.rows()),
                "Number of alpha parameters must be equal to number of rows of X");
    BNMF_ASSERT(p.beta.size() == z, "Number of beta parameters must be equal to z");
}

template <typename T, typename Scalar>
tensor_t<T, 3> solve(const matrix_t<T>& X, size_t z, const alloc_model::Params<Scalar>& p,
                     const T* S0, size_t max_iter, bool use_psi_appr, double eps, uint64_t seed) {
    check_bld_params(X, z, p);
    const size_t x = static_cast<size_t>(X.rows()), y = static_cast<size_t>(X.cols());
    tensor_t<T, 3> S(static_cast<long>(x), static_cast<long>(y), static_cast<long>(z));
    const std::vector<double> alpha = bnmf_algs::details::to_double(p.alpha);
    const std::vector<double> beta = bnmf_algs::details::to_double(p.beta);
    bnmf_status st;
    if constexpr (bnmf_algs::details::dtype_of<T>() == BNMF_F64)
        st = bnmf_bld_mult_f64(X.data(), x, y, z, alpha.data(), beta.data(), max_iter,
                               use_psi_appr ? 1 : 0, eps, S0, seed, S.data());
    else
        st = bnmf_bld_mult_f32(X.data(), x, y, z, alpha.data(), beta.data(), max_iter,
                               use_psi_appr ? 1 : 0, eps, S0, seed, S.data());
    bnmf_algs::details::check_status(st, "bld_mult");
    return S;
}
}  // namespace details

template <typename T, typename Scalar>
tensor_t<T, 3> bld_mult(const matrix_t<T>& X, const size_t z,
                        const alloc_model::Params<Scalar>& model_params, size_t max_iter = 1000,
                        bool use_psi_appr = false, double eps = 1e-50) {
    return details::solve<T, Scalar>(X, z, model_params, nullptr, max_iter, use_psi_appr, eps,
                                     kDefaultInitSeed);
}

// Starts from the caller's S0 (x*y*z) instead of a drawn one.
template <typename T, typename Scalar>
tensor_t<T, 3> bld_mult(const matrix_t<T>& X, const size_t z,
                        const alloc_model::Params<Scalar>& model_params, const tensor_t<T, 3>& S0,
                        size_t max_iter = 1000, bool use_psi_appr = false, double eps = 1e-50) {
    BNMF_ASSERT(S0.dimension(0) == X.rows() && S0.dimension(1) == X.cols() &&
                    static_cast<size_t>(S0.dimension(2)) == z,
                "S0 must have shape (x, y, z)");
    return details::solve<T, Scalar>(X, z, model_params, S0.data(), max_iter, use_psi_appr, eps,
                                     kDefaultInitSeed);
}

// The reference's GPU driver has the same contract; both entry points run on
// the GPU here.
template <typename T, typename Scalar>
tensor_t<T, 3> bld_mult_cuda(const matrix_t<T>& X, const size_t z,
                             const alloc_model::Params<Scalar>& model_params,
                             size_t max_iter = 1000, bool use_psi_appr = false,
                             double eps = 1e-50) {
    return bld_mult<T, Scalar>(X, z, model_params, max_iter, use_psi_appr, eps);
}

template <typename T, typename Scalar>
tensor_t<T, 3> bld_mult_cuda(const matrix_t<T>& X, const size_t z,
                             const alloc_model::Params<Scalar>& model_params,
                             const tensor_t<T, 3>& S0, size_t max_iter = 1000,
                             bool use_psi_appr = false, double eps = 1e-50) {
    return bld_mult<T, Scalar>(X, z, model_params, S0, max_iter, use_psi_appr, eps);
}

}  // namespace bld
}  // namespace bnmf_algs
