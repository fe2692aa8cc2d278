// defs.hpp — types of the drop-in C++ API (role of src/defs.hpp:26-66).
//
// With Eigen3 (and its unsupported Tensor module) on the include path the
// aliases are the reference's own: Eigen RowMajor matrices and tensors, so
// bld::bld_mult has exactly the reference signature. Without Eigen a small
// RowMajor container with the same alias names and the subset of the Eigen
// API used by bld_mult callers (rows/cols/dimension/data/operator(),
// Constant/Zero, comma initialisation, isApprox) stands in; define
// BNMF_HIP_NO_EIGEN to force it.
#pragma once

#include <array>
#include <cmath>
#include <cstddef>
#include <initializer_list>
#include <stdexcept>
#include <vector>

#if !defined(BNMF_HIP_NO_EIGEN) && __has_include(<Eigen/Dense>) && \
    __has_include(<unsupported/Eigen/CXX11/Tensor>)
#define BNMF_HIP_EIGEN 1
#include <Eigen/Dense>
#include <unsupported/Eigen/CXX11/Tensor>
#endif

// Parameter errors throw std::invalid_argument, as the reference's
// BNMF_ASSERT does when it is built for testing (src/defs.hpp:12-16).
#define BNMF_ASSERT(condition, msg) \
    if (!(condition)) throw std::invalid_argument(msg)

namespace bnmf_algs {

#ifdef BNMF_HIP_EIGEN
template <typename Scalar>
using vector_t = Eigen::Matrix<Scalar, 1, Eigen::Dynamic, Eigen::RowMajor>;
template <typename Scalar>
using matrix_t = Eigen::Matrix<Scalar, Eigen::Dynamic, Eigen::Dynamic, Eigen::RowMajor>;
template <typename Scalar, size_t N>
using tensor_t = Eigen::Tensor<Scalar, N, Eigen::RowMajor>;
template <size_t N>
using shape = Eigen::array<size_t, N>;
#else
namespace shim {

// Dense row-major matrix (a vector is a 1 x n matrix).
template <typename Scalar>
class Matrix {
   public:
    Matrix() = default;
    Matrix(long rows, long cols) : r_(rows), c_(cols), d_((size_t)(rows * cols), Scalar(0)) {}
    long rows() const { return r_; }
    long cols() const { return c_; }
    long size() const { return r_ * c_; }
    Scalar* data() { return d_.data(); }
    const Scalar* data() const { return d_.data(); }
    Scalar& operator()(long i, long j) { return d_[(size_t)(i * c_ + j)]; }
    const Scalar& operator()(long i, long j) const { return d_[(size_t)(i * c_ + j)]; }
    Scalar& operator()(long i) { return d_[(size_t)i]; }
    const Scalar& operator()(long i) const { return d_[(size_t)i]; }
    static Matrix Constant(long r, long c, Scalar v) {
        Matrix m(r, c);
        for (auto& e : m.d_) e = v;
        return m;
    }
    static Matrix Zero(long r, long c) { return Matrix(r, c); }
    // comma initialiser: m << a, b, c, ...;
    struct CommaInit {
        Matrix* m;
        size_t pos;
        CommaInit& operator,(Scalar v) {
            if (pos >= m->d_.size()) throw std::out_of_range("too many coefficients");
            m->d_[pos++] = v;
            return *this;
        }
    };
    CommaInit operator<<(Scalar v) {
        if (d_.empty()) throw std::out_of_range("too many coefficients");
        d_[0] = v;
        return CommaInit{this, 1};
    }
    // Eigen's isApprox: ||a - b|| <= prec * min(||a||, ||b||)
    bool isApprox(const Matrix& o, double prec = 1e-12) const {
        double d = 0, na = 0, nb = 0;
        for (size_t t = 0; t < d_.size(); ++t) {
            const double a = (double)d_[t], b = (double)o.d_[t];
            d += (a - b) * (a - b);
            na += a * a;
            nb += b * b;
        }
        return std::sqrt(d) <= prec * std::sqrt(std::min(na, nb));
    }

   private:
    long r_ = 0, c_ = 0;
    std::vector<Scalar> d_;
};

// Dense row-major N-dimensional tensor (last index fastest).
template <typename Scalar, size_t N>
class Tensor {
   public:
    Tensor() { dims_.fill(0); }
    template <typename... D>
    explicit Tensor(D... dims) : dims_{{(long)dims...}} {
        static_assert(sizeof...(D) == N, "one extent per dimension");
        size_t n = 1;
        for (long d : dims_) n *= (size_t)d;
        d_.assign(n, Scalar(0));
    }
    long dimension(size_t i) const { return dims_[i]; }
    long size() const { return (long)d_.size(); }
    Scalar* data() { return d_.data(); }
    const Scalar* data() const { return d_.data(); }
    template <typename... I>
    Scalar& operator()(I... idx) {
        return d_[offset({(long)idx...})];
    }
    template <typename... I>
    const Scalar& operator()(I... idx) const {
        return d_[offset({(long)idx...})];
    }
    void setZero() { setConstant(Scalar(0)); }
    void setConstant(Scalar v) {
        for (auto& e : d_) e = v;
    }

   private:
    size_t offset(std::initializer_list<long> idx) const {
        size_t o = 0, n = 0;
        for (long i : idx) o = o * (size_t)dims_[n++] + (size_t)i;
        return o;
    }
    std::array<long, N> dims_;
    std::vector<Scalar> d_;
};

}  // namespace shim

template <typename Scalar>
using vector_t = shim::Matrix<Scalar>;
template <typename Scalar>
using matrix_t = shim::Matrix<Scalar>;
template <typename Scalar, size_t N>
using tensor_t = shim::Tensor<Scalar, N>;
template <size_t N>
using shape = std::array<size_t, N>;
#endif

using vectord = vector_t<double>;
using matrixd = matrix_t<double>;
template <size_t N>
using tensord = tensor_t<double, N>;

}  // namespace bnmf_algs
