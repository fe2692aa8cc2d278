// alloc_model_funcs.hpp — log_marginal_S on the GPU (role of
// src/alloc_model/alloc_model_funcs.hpp:574-586) and seeded input synthesis
// (the laws of bnmf_priors / sample_S, :463-552).
#pragma once

#include <cstdint>

#include "../defs.hpp"
#include "../detail.hpp"
#include "alloc_model_params.hpp"

namespace bnmf_algs {
namespace alloc_model {

template <typename T, typename Scalar>
double log_marginal_S(const tensor_t<T, 3>& S, const Params<Scalar>& model_params) {
    BNMF_ASSERT(model_params.alpha.size() == static_cast<size_t>(S.dimension(0)),
                "Number of alpha parameters must be equal to S.dimension(0)");
    BNMF_ASSERT(model_params.beta.size() == static_cast<size_t>(S.dimension(2)),
                "Number of alpha parameters must be equal to S.dimension(2)");
    const std::vector<double> alpha = bnmf_algs::details::to_double(model_params.alpha);
    const std::vector<double> beta = bnmf_algs::details::to_double(model_params.beta);
    double out = 0.0;
    bnmf_algs::details::check_status(
        bnmf_log_marginal_S(bnmf_algs::details::dtype_of<T>(), S.data(),
                            static_cast<size_t>(S.dimension(0)), static_cast<size_t>(S.dimension(1)),
                            static_cast<size_t>(S.dimension(2)), alpha.data(), beta.data(),
                            static_cast<double>(model_params.a), static_cast<double>(model_params.b),
                            &out),
        "log_marginal_S");
    return out;
}

// X = S_ij+ of an allocation-model sample drawn from `seed`.
template <typename Scalar>
matrix_t<double> sample_X(uint64_t seed, const shape<3>& s, const Params<Scalar>& p) {
    const std::vector<double> alpha = bnmf_algs::details::to_double(p.alpha);
    const std::vector<double> beta = bnmf_algs::details::to_double(p.beta);
    matrix_t<double> X(static_cast<long>(s[0]), static_cast<long>(s[1]));
    bnmf_algs::details::check_status(bnmf_sample_X(seed, s[0], s[1], s[2], alpha.data(), beta.data(),
                                                   (double)p.a, (double)p.b, X.data()),
                                     "sample_X");
    return X;
}

// A full allocation tensor S drawn from `seed` (sample_S after bnmf_priors).
template <typename Scalar>
tensor_t<double, 3> sample_S(uint64_t seed, const shape<3>& s, const Params<Scalar>& p) {
    const std::vector<double> alpha = bnmf_algs::details::to_double(p.alpha);
    const std::vector<double> beta = bnmf_algs::details::to_double(p.beta);
    tensor_t<double, 3> S(static_cast<long>(s[0]), static_cast<long>(s[1]), static_cast<long>(s[2]));
    bnmf_algs::details::check_status(bnmf_sample_S(seed, s[0], s[1], s[2], alpha.data(), beta.data(),
                                                   (double)p.a, (double)p.b, S.data()),
                                     "sample_S");
    return S;
}

// init_S (bld_mult_funcs.hpp:36-74) from `seed`.
template <typename T>
tensor_t<T, 3> init_S(uint64_t seed, const matrix_t<T>& X, size_t z) {
    tensor_t<T, 3> S(X.rows(), X.cols(), static_cast<long>(z));
    bnmf_algs::details::check_status(
        bnmf_init_S(bnmf_algs::details::dtype_of<T>(), seed, X.data(), static_cast<size_t>(X.rows()),
                    static_cast<size_t>(X.cols()), z, S.data()),
        "init_S");
    return S;
}

}  // namespace alloc_model
}  // namespace bnmf_algs
