// alloc_model_params.hpp — allocation-model parameters
// (role of src/alloc_model/alloc_model_params.hpp:25-99).
#pragma once

#include <vector>

#include "../defs.hpp"

namespace bnmf_algs {
namespace alloc_model {

// L_j ~ Gamma(a, b), W_:k ~ Dirichlet(alpha) (length x),
// H_:j ~ Dirichlet(beta) (length z); S_ijk ~ Poisson(W_ik H_kj L_j).
template <typename Scalar>
struct Params {
    Scalar a;
    Scalar b;
    std::vector<Scalar> alpha;
    std::vector<Scalar> beta;

    Params(Scalar a, Scalar b, const std::vector<Scalar>& alpha, const std::vector<Scalar>& beta)
        : a(a), b(b), alpha(alpha), beta(beta) {}
    // a = 1, b = 10, alpha = 1 (x), beta = 1 (z)
    explicit Params(const shape<3>& tensor_shape)
        : a(static_cast<Scalar>(1)), b(static_cast<Scalar>(10)),
          alpha(tensor_shape[0], static_cast<Scalar>(1)),
          beta(tensor_shape[2], static_cast<Scalar>(1)) {}
    Params() = default;
};

template <typename Scalar>
Params<Scalar> make_bld_params(const shape<3>& tensor_shape) {
    return Params<Scalar>(tensor_shape);
}

}  // namespace alloc_model
}  // namespace bnmf_algs
