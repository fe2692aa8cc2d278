// sampler.cpp — seeded input synthesis for the BLD solver (host code).
//
// Distributions follow the reference's allocation model
// (src/alloc_model/alloc_model_funcs.hpp:463-552) and init_S
// (src/bld/bld_mult/bld_mult_funcs.hpp:36-74); the reference draws from GSL's
// taus generator, whose stream is not reproduced. Here every draw comes from
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11) keyed by the seed, with the counter naming the
// variate, so results are identical for any thread count.
//   Gamma(a >= 1): Marsaglia & Tsang (2000); a < 1: Gamma(a+1) * U^(1/a).
//   Normal: Box-Muller.  Dirichlet: normalised Gammas.
//   Poisson: Knuth's product method for mean < 12, otherwise the
//   transformed rejection method PTRS (Hormann 1993) (poisson_draw,
//   philox.hpp, shared with the device sampler).
// gsl_ran_gamma(r, a, b) takes b as the SCALE; the same convention is kept.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "bnmf_internal.hpp"
#include "philox.hpp"

namespace bnmf {
namespace {

double normal_draw(Philox& g) {  // Box-Muller
    const double u1 = g.uniform(), u2 = g.uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

double gamma_draw(Philox& g, double shape, double scale) {
    if (shape < 1.0) {
        const double u = g.uniform();
        return gamma_draw(g, shape + 1.0, scale) * std::pow(u, 1.0 / shape);
    }
    const double d = shape - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
    for (;;) {
        double xn, v;
        do {
            xn = normal_draw(g);
            v = 1.0 + c * xn;
        } while (v <= 0.0);
        v = v * v * v;
        const double u = g.uniform();
        if (u < 1.0 - 0.0331 * xn * xn * xn * xn) return scale * d * v;
        if (std::log(u) < 0.5 * xn * xn + d * (1.0 - v + std::log(v))) return scale * d * v;
    }
}

template <typename F>
void parallel_for(size_t n, F&& f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096 || nt == 1) {
        for (size_t t = 0; t < n; ++t) f(t);
        return;
    }
    std::vector<std::thread> th;
    const size_t chunk = (n + nt - 1) / nt;
    for (unsigned w = 0; w < nt; ++w) {
        const size_t b = w * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([b, e, &f] {
            for (size_t t = b; t < e; ++t) f(t);
        });
    }
    for (auto& t : th) t.join();
}

// priors of bnmf_priors (alloc_model_funcs.hpp:463-505)
void priors(uint64_t seed, size_t x, size_t y, size_t z, const double* alpha,
            const double* beta, double a, double b, std::vector<double>& W,
            std::vector<double>& H, std::vector<double>& L) {
    L.resize(y);
    for (size_t j = 0; j < y; ++j) {
        Philox g(seed, ST_L, j);
        L[j] = gamma_draw(g, a, b);
    }
    W.assign(x * z, 0.0);  // x*z, columns ~ Dir(alpha)
    parallel_for(z, [&](size_t k) {
        Philox g(seed, ST_W, k);
        double s = 0.0;
        std::vector<double> v(x);
        for (size_t i = 0; i < x; ++i) {
            v[i] = gamma_draw(g, alpha[i], 1.0);
            s += v[i];
        }
        for (size_t i = 0; i < x; ++i) W[i * z + k] = v[i] / s;
    });
    H.assign(z * y, 0.0);  // z*y, columns ~ Dir(beta)
    parallel_for(y, [&](size_t j) {
        Philox g(seed, ST_H, j);
        double s = 0.0;
        std::vector<double> v(z);
        for (size_t k = 0; k < z; ++k) {
            v[k] = gamma_draw(g, beta[k], 1.0);
            s += v[k];
        }
        for (size_t k = 0; k < z; ++k) H[k * y + j] = v[k] / s;
    });
}

}  // namespace

void sample_priors(uint64_t seed, size_t x, size_t y, size_t z, const double* alpha,
                   const double* beta, double a, double b, std::vector<double>& W,
                   std::vector<double>& H, std::vector<double>& L) {
    priors(seed, x, y, z, alpha, beta, a, b, W, H, L);
}

int sample_X(uint64_t seed, size_t x, size_t y, size_t z, const double* alpha,
             const double* beta, double a, double b, double* X) {
    std::vector<double> W, H, L;
    priors(seed, x, y, z, alpha, beta, a, b, W, H, L);
    parallel_for(x, [&](size_t i) {
        for (size_t j = 0; j < y; ++j) {
            double lam = 0.0;
            for (size_t k = 0; k < z; ++k) lam += W[i * z + k] * H[k * y + j];
            Philox g(seed, ST_X, (uint64_t)i * y + j);
            X[i * y + j] = poisson_draw(g, lam * L[j]);
        }
    });
    return 0;
}

int sample_S(uint64_t seed, size_t x, size_t y, size_t z, const double* alpha,
             const double* beta, double a, double b, double* S) {
    std::vector<double> W, H, L;
    priors(seed, x, y, z, alpha, beta, a, b, W, H, L);
    parallel_for(x, [&](size_t i) {
        for (size_t j = 0; j < y; ++j)
            for (size_t k = 0; k < z; ++k) {
                Philox g(seed, ST_S, ((uint64_t)i * y + j) * z + k);
                S[(i * y + j) * z + k] = poisson_draw(g, W[i * z + k] * H[k * y + j] * L[j]);
            }
    });
    return 0;
}

template <typename T>
int init_S(uint64_t seed, const T* X, size_t x, size_t y, size_t z, T* S) {
    parallel_for(x * y, [&](size_t f) {
        thread_local std::vector<double> w;
        if (w.size() < z) w.resize(z);
        init_fiber<T>(seed, f, (double)X[f], (uint32_t)z, S + f * z, w.data());
    });
    return 0;
}

template int init_S<float>(uint64_t, const float*, size_t, size_t, size_t, float*);
template int init_S<double>(uint64_t, const double*, size_t, size_t, size_t, double*);

}  // namespace bnmf
