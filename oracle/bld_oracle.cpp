// bld_oracle.cpp — FP64 CPU oracle for the multiplicative BLD solver.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker, never the product:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// the library built from it (oracle/liboracle.so). Nothing in the shipped
// library links, includes or calls it.
//
// It restates, step for step and in double precision, the reference CPU
// solver of eozd/bnmf-algs (paths relative to the reference root):
//   bld::bld_mult                      src/bld/bld_mult/bld_mult.hpp:83-150
//   details::check_bld_params          src/bld/util_details.hpp:20-28
//   details::bld_mult::X_reciprocal    src/bld/bld_mult/bld_mult_funcs.hpp:87-100
//   update_alpha_eph / update_beta_eph bld_mult_funcs.hpp:130-141 / 152-163
//   update_grad_plus                   bld_mult_funcs.hpp:177-192
//   update_grad_minus                  bld_mult_funcs.hpp:205-229
//   update_nom_mult / update_denom_mult bld_mult_funcs.hpp:242-259 / 272-290
//   update_S                           bld_mult_funcs.hpp:307-327
//   util::psi_appr                     src/util/util.hpp:440-514
//   util::sparseness                   src/util/util.hpp:270-295
//   alloc_model::log_marginal_S        src/alloc_model/alloc_model_funcs.hpp:574-586
//     (terms: first :34-77, second :91-132, third :146-169, fourth :179-194)
//   bld::bld_fact                      src/bld/bld_fact.hpp:36-80
//   bld::bld_add                       src/bld/bld_add.hpp:69-184
//   bld::bld_appr                      src/bld/bld_appr.hpp:1-372
//   util::normalize (L1, along k)      src/util/util.hpp:321-383
//
// The reference's exact digamma is GSL's gsl_sf_psi (util_details.hpp:71-73),
// which is absent from this machine. oracle_psi below restates the digamma
// from published mathematics: the recurrence psi(x) = psi(x+1) - 1/x
// (Abramowitz & Stegun 6.3.5) lifts x to >= 10, then the asymptotic series
// psi(x) ~ ln x - 1/(2x) - sum_n B_2n / (2n x^2n) (A&S 6.3.18) is summed to
// the x^-16 term. lnGamma (gsl_sf_lngamma in the reference) is std::lgamma.
//
// Parity status of each function (how it is pinned) is listed in
// tests/test_oracle.py; every function here is checked against mpmath golden
// vectors, closed forms or brute force on tiny inputs.
//
// Determinism: every reduction runs in a fixed sequential order, and OpenMP
// only splits independent output elements, so results do not depend on the
// thread count.

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;

// Exact digamma for real x (role of gsl_sf_psi).
double psi_exact(double x) {
    if (std::isnan(x)) return x;
    if (x <= 0.0) {
        if (x == std::floor(x)) return std::numeric_limits<double>::quiet_NaN();
        // reflection: psi(1-x) - psi(x) = pi cot(pi x)   (A&S 6.3.7)
        return psi_exact(1.0 - x) - kPi / std::tan(kPi * x);
    }
    double acc = 0.0;
    while (x < 10.0) {  // A&S 6.3.5, one exact reciprocal per step
        acc -= 1.0 / x;
        x += 1.0;
    }
    const double r = 1.0 / x;
    const double r2 = r * r;
    // B2/2, -B4/4, B6/6, ... (Bernoulli numbers B2..B16)
    const double series =
        r2 * (1.0 / 12 -
              r2 * (1.0 / 120 -
                    r2 * (1.0 / 252 -
                          r2 * (1.0 / 240 -
                                r2 * (1.0 / 132 -
                                      r2 * (691.0 / 32760 -
                                            r2 * (1.0 / 12 -
                                                  r2 * (3617.0 / 8160))))))));
    return acc + (std::log(x) - 0.5 * r - series);
}

// util::psi_appr (util.hpp:440-514): shift x below 6 up by the recurrence,
// case by case as the reference writes it, then a closed-form series to x^-14.
// The operation order follows the reference's so that double results agree.
double psi_appr(double x) {
    double extra = 0;
    if (x < 1) {
        const double a = x + 1, b = x + 2, c = x + 3, d = x + 4, e = x + 5;
        const double ab = a * b, cd = c * d, ex = e * x;
        extra = ((a + b) * cd * ex + ab * d * ex + ab * c * ex + ab * cd * x +
                 ab * cd * e) /
                (ab * cd * ex);
        x += 6;
    } else if (x < 2) {
        const double a = x + 1, b = x + 2, c = x + 3, d = x + 4;
        const double ab = a * b, cd = c * d, dx = d * x;
        extra = ((a + b) * c * dx + ab * dx + ab * c * x + ab * cd) / (ab * cd * x);
        x += 5;
    } else if (x < 3) {
        const double a = x + 1, b = x + 2, c = x + 3;
        const double ab = a * b, cx = c * x;
        extra = ((a + b) * cx + (c + x) * ab) / (ab * cx);
        x += 4;
    } else if (x < 4) {
        const double a = x + 1, b = x + 2;
        const double ab = a * b;
        extra = ((a + b) * x + ab) / (ab * x);
        x += 3;
    } else if (x < 5) {
        const double a = x + 1;
        extra = (a + x) / (a * x);
        x += 2;
    } else if (x < 6) {
        extra = 1 / x;
        x += 1;
    }
    const double x2 = x * x, x4 = x2 * x2, x6 = x4 * x2, x8 = x6 * x2;
    const double x10 = x8 * x2, x12 = x10 * x2, x13 = x12 * x, x14 = x13 * x;
    const double res =
        std::log(x) + (-360360 * x13 - 60060 * x12 + 6006 * x10 - 2860 * x8 +
                       3003 * x6 - 5460 * x4 + 15202 * x2 - 60060) /
                          (720720 * x14);
    return res - extra;
}

inline double psi_sel(double x, int appr) { return appr ? psi_appr(x) : psi_exact(x); }

struct Dims {
    size_t x, y, z;
    size_t at(size_t i, size_t j, size_t k) const { return (i * y + j) * z + k; }
};

void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

// Axis sums in sequential order (bld_mult.hpp:124-133).
void tensor_sums(const double* S, const Dims& d, double* S_pjk, double* S_ipk,
                 double* S_ijp) {
    const long x = (long)d.x, y = (long)d.y;
    const size_t z = d.z;
    if (S_pjk) {
#pragma omp parallel for schedule(static)
        for (long j = 0; j < y; ++j) {
            double* out = S_pjk + (size_t)j * z;
            for (size_t k = 0; k < z; ++k) out[k] = 0.0;
            for (long i = 0; i < x; ++i) {
                const double* f = S + d.at(i, j, 0);
                for (size_t k = 0; k < z; ++k) out[k] += f[k];
            }
        }
    }
    if (S_ipk) {
#pragma omp parallel for schedule(static)
        for (long i = 0; i < x; ++i) {
            double* out = S_ipk + (size_t)i * z;
            for (size_t k = 0; k < z; ++k) out[k] = 0.0;
            for (long j = 0; j < y; ++j) {
                const double* f = S + d.at(i, j, 0);
                for (size_t k = 0; k < z; ++k) out[k] += f[k];
            }
        }
    }
    if (S_ijp) {
#pragma omp parallel for schedule(static)
        for (long i = 0; i < x; ++i) {
            for (long j = 0; j < y; ++j) {
                const double* f = S + d.at(i, j, 0);
                double s = 0.0;
                for (size_t k = 0; k < z; ++k) s += f[k];
                S_ijp[(size_t)i * d.y + j] = s;
            }
        }
    }
}

// Intermediates of one iteration, all row-major.
struct Work {
    std::vector<double> S_pjk, S_ipk, S_ijp, alpha_eph, beta_eph, psi_beta_eph,
        grad_minus, nom, denom, X_recip;
    // grad_plus is only materialised when a trace is requested.
    std::vector<double> grad_plus;
};

// One iteration of bld_mult.hpp:123-147, in place on S.
void iterate(const double* X, const Dims& d, const double* alpha,
             const double* beta, int appr, double eps, double* S, Work& w,
             bool keep_grad_plus) {
    const long x = (long)d.x, y = (long)d.y;
    const size_t z = d.z;
    tensor_sums(S, d, w.S_pjk.data(), w.S_ipk.data(), w.S_ijp.data());

    // update_alpha_eph (funcs:130-141), update_beta_eph (funcs:152-163)
    for (long i = 0; i < x; ++i)
        for (size_t k = 0; k < z; ++k)
            w.alpha_eph[(size_t)i * z + k] = w.S_ipk[(size_t)i * z + k] + alpha[i];
    for (long j = 0; j < y; ++j)
        for (size_t k = 0; k < z; ++k)
            w.beta_eph[(size_t)j * z + k] = w.S_pjk[(size_t)j * z + k] + beta[k];

    // update_grad_plus (funcs:177-192) evaluates psi(beta_eph(j,k)) for every
    // i; psi is a pure function, so evaluating it once per (j,k) gives the
    // identical value.
#pragma omp parallel for schedule(static)
    for (long jk = 0; jk < y * (long)z; ++jk)
        w.psi_beta_eph[jk] = psi_sel(w.beta_eph[jk], appr);

    // update_grad_minus (funcs:205-229)
    std::vector<double> psi_alpha_eph_sum(z);
    for (size_t k = 0; k < z; ++k) {
        double sum = 0.0;
        for (long i = 0; i < x; ++i) sum += w.alpha_eph[(size_t)i * z + k];
        psi_alpha_eph_sum[k] = psi_sel(sum, appr);
    }
    for (long i = 0; i < x; ++i)
        for (size_t k = 0; k < z; ++k)
            w.grad_minus[(size_t)i * z + k] =
                psi_alpha_eph_sum[k] - psi_sel(w.alpha_eph[(size_t)i * z + k], appr);

    // Fused per-fiber: grad_plus, nom_mult (funcs:242-259),
    // denom_mult (funcs:272-290), update_S (funcs:307-327). Every fiber is
    // independent; within a fiber the k order is the reference's.
#pragma omp parallel
    {
        std::vector<double> gp(z);
#pragma omp for schedule(static)
        for (long i = 0; i < x; ++i) {
            const double* gm = w.grad_minus.data() + (size_t)i * z;
            for (long j = 0; j < y; ++j) {
                double* s = S + d.at(i, j, 0);
                const double* pb = w.psi_beta_eph.data() + (size_t)j * z;
                const double xr = w.X_recip[(size_t)i * d.y + j];
                for (size_t k = 0; k < z; ++k) gp[k] = pb[k] - psi_sel(s[k] + 1, appr);
                double nom_sum = 0.0, denom_sum = 0.0;
                for (size_t k = 0; k < z; ++k) nom_sum += (s[k] * xr * gm[k]);
                for (size_t k = 0; k < z; ++k) denom_sum += (s[k] * xr * gp[k]);
                w.nom[(size_t)i * d.y + j] = nom_sum;
                w.denom[(size_t)i * d.y + j] = denom_sum;
                if (keep_grad_plus)
                    std::memcpy(w.grad_plus.data() + d.at(i, j, 0), gp.data(), z * sizeof(double));
                const double s_ij = w.S_ijp[(size_t)i * d.y + j];
                const double x_over_s = X[(size_t)i * d.y + j] / (s_ij + eps);
                for (size_t k = 0; k < z; ++k)
                    s[k] *= (gp[k] + nom_sum) * x_over_s / (gm[k] + denom_sum + eps);
            }
        }
    }
}

// One iteration of bld::bld_add (src/bld/bld_add.hpp:108-181), in place on
// S, at iteration number eph (the step size depends on it). Per fiber, in the
// reference's operation order:
//   grad_k = psi_appr(beta_eph_jk) - psi_appr(S_k + 1)
//            - psi_appr(sum_i alpha_eph_ik) + psi_appr(alpha_eph_ik)  (:137-146)
//   S_mult = sum_k S_k grad_k                                         (:148-155)
//   Z_k    = (log S_k + eps) + eta(eph) S_k (X grad_k - S_mult)       (:150,158-166)
//   Z_k    = exp(Z_k) / (sum_k |exp(Z_k)| + 1e-50)   util::normalize's default
//            eps (util.hpp:321-383), not the solver's eps           (:167-173)
//   S_k    = X Z_k                                                    (:175-181)
// eta(step) = 0.1 / (step + 1)^0.55 (:101-103). Every psi is psi_appr.
// With grad_out non-null, grad_S of the iteration is stored there.
void iterate_add(const double* X, const Dims& d, const double* alpha, const double* beta,
                 double eps, size_t eph, double* S, Work& w, double* grad_out) {
    const long x = (long)d.x, y = (long)d.y;
    const size_t z = d.z;
    tensor_sums(S, d, w.S_pjk.data(), w.S_ipk.data(), nullptr);
    for (long i = 0; i < x; ++i)
        for (size_t k = 0; k < z; ++k)
            w.alpha_eph[(size_t)i * z + k] = alpha[i] + w.S_ipk[(size_t)i * z + k];
    for (long j = 0; j < y; ++j)
        for (size_t k = 0; k < z; ++k)
            w.beta_eph[(size_t)j * z + k] = beta[k] + w.S_pjk[(size_t)j * z + k];
    std::vector<double> alpha_eph_sum(z, 0.0);
    for (long i = 0; i < x; ++i)
        for (size_t k = 0; k < z; ++k) alpha_eph_sum[k] += w.alpha_eph[(size_t)i * z + k];
    const double eta = 0.1 / std::pow((double)(eph + 1), 0.55);
#pragma omp parallel
    {
        std::vector<double> g(z), e(z);
#pragma omp for schedule(static)
        for (long i = 0; i < x; ++i) {
            for (long j = 0; j < y; ++j) {
                double* s = S + d.at(i, j, 0);
                const double xv = X[(size_t)i * d.y + j];
                for (size_t k = 0; k < z; ++k)
                    g[k] = psi_appr(w.beta_eph[(size_t)j * z + k]) - psi_appr(s[k] + 1) -
                           psi_appr(alpha_eph_sum[k]) + psi_appr(w.alpha_eph[(size_t)i * z + k]);
                double s_mult = 0.0;
                for (size_t k = 0; k < z; ++k) s_mult += s[k] * g[k];
                double norm = 0.0;
                for (size_t k = 0; k < z; ++k) {
                    double zk = std::log(s[k]) + eps;
                    zk += eta * s[k] * (xv * g[k] - s_mult);
                    e[k] = std::exp(zk);
                    norm += std::fabs(e[k]);
                }
                norm += 1e-50;
                if (grad_out) std::memcpy(grad_out + d.at(i, j, 0), g.data(), z * sizeof(double));
                for (size_t k = 0; k < z; ++k) s[k] = xv * (e[k] / norm);
            }
        }
    }
}

// bld::bld_appr (src/bld/bld_appr.hpp:287-372) state: nu (x*z, rows sum to
// 1) and mu (z*y, columns sum to 1); S = (X / (nu mu + eps)) nu_ik mu_kj.
struct Appr {
    std::vector<double> X_hat, oa, oa2, S, S_pjk, S_ipk, alpha_eph, beta_eph, gp, gm;
};

// The helper updates of bld_appr.hpp:322-330 (update_X_hat :12-15,
// update_orig_over_appr :17-23, ..._squared :25-32, update_S :34-48,
// update_alpha_eph :50-60, update_beta_eph :62-71, update_grad_plus :73-88,
// update_grad_minus :90-103). Every psi is psi_appr.
void appr_helpers(const double* X, const Dims& d, const double* alpha, const double* beta,
                  double eps, const double* nu, const double* mu, Appr& a) {
    const size_t x = d.x, y = d.y, z = d.z;
    a.X_hat.assign(x * y, 0.0);
    a.oa.assign(x * y, 0.0);
    a.oa2.assign(x * y, 0.0);
    a.S.assign(x * y * z, 0.0);
    a.S_pjk.assign(y * z, 0.0);
    a.S_ipk.assign(x * z, 0.0);
    a.alpha_eph.assign(x * z, 0.0);
    a.beta_eph.assign(y * z, 0.0);
    a.gp.assign(x * y * z, 0.0);
    a.gm.assign(x * z, 0.0);
    for (size_t i = 0; i < x; ++i)
        for (size_t j = 0; j < y; ++j) {
            double h = 0.0;
            for (size_t k = 0; k < z; ++k) h += nu[i * z + k] * mu[k * y + j];
            a.X_hat[i * y + j] = h;
            a.oa[i * y + j] = X[i * y + j] / (h + eps);
            a.oa2[i * y + j] = X[i * y + j] / (std::pow(h, 2) + eps);
        }
    for (size_t i = 0; i < x; ++i)
        for (size_t j = 0; j < y; ++j)
            for (size_t k = 0; k < z; ++k)
                a.S[d.at(i, j, k)] = a.oa[i * y + j] * nu[i * z + k] * mu[k * y + j];
    tensor_sums(a.S.data(), d, a.S_pjk.data(), a.S_ipk.data(), nullptr);
    for (size_t i = 0; i < x; ++i)
        for (size_t k = 0; k < z; ++k) a.alpha_eph[i * z + k] = a.S_ipk[i * z + k] + alpha[i];
    for (size_t j = 0; j < y; ++j)
        for (size_t k = 0; k < z; ++k) a.beta_eph[j * z + k] = a.S_pjk[j * z + k] + beta[k];
    for (size_t i = 0; i < x; ++i)
        for (size_t j = 0; j < y; ++j)
            for (size_t k = 0; k < z; ++k)
                a.gp[d.at(i, j, k)] = psi_appr(a.beta_eph[j * z + k]) - psi_appr(a.S[d.at(i, j, k)] + 1);
    std::vector<double> A(z, 0.0);
    for (size_t i = 0; i < x; ++i)
        for (size_t k = 0; k < z; ++k) A[k] += a.alpha_eph[i * z + k];
    for (size_t i = 0; i < x; ++i)
        for (size_t k = 0; k < z; ++k) a.gm[i * z + k] = psi_appr(A[k]) - psi_appr(a.alpha_eph[i * z + k]);
}

// update_nu (bld_appr.hpp:105-164): the four accumulations in the
// reference's loop order, the multiplicative step and the row
// normalisation nu_i: /= (sum_k nu_ik + eps).
void appr_update_nu(const Dims& d, double eps, const Appr& a, const double* mu, double* nu,
                    double* nom_out = nullptr, double* den_out = nullptr) {
    const size_t x = d.x, y = d.y, z = d.z;
    std::vector<double> nom(x * z, 0.0), den(x * z, 0.0);
    for (size_t i = 0; i < x; ++i)
        for (size_t j = 0; j < y; ++j)
            for (size_t k = 0; k < z; ++k)
                nom[i * z + k] += a.oa[i * y + j] * mu[k * y + j] * a.gp[d.at(i, j, k)];
    for (size_t i = 0; i < x; ++i)
        for (size_t j = 0; j < y; ++j)
            for (size_t k = 0; k < z; ++k)
                for (size_t c = 0; c < z; ++c)
                    nom[i * z + k] += a.oa2[i * y + j] * mu[k * y + j] * nu[i * z + c] * mu[c * y + j] *
                                      a.gm[i * z + c];
    for (size_t i = 0; i < x; ++i)
        for (size_t j = 0; j < y; ++j)
            for (size_t k = 0; k < z; ++k)
                den[i * z + k] += a.oa[i * y + j] * mu[k * y + j] * a.gm[i * z + k];
    for (size_t i = 0; i < x; ++i)
        for (size_t j = 0; j < y; ++j)
            for (size_t k = 0; k < z; ++k)
                for (size_t c = 0; c < z; ++c)
                    den[i * z + k] += a.oa2[i * y + j] * mu[k * y + j] * nu[i * z + c] * mu[c * y + j] *
                                      a.gp[d.at(i, j, c)];
    if (nom_out) std::memcpy(nom_out, nom.data(), x * z * sizeof(double));
    if (den_out) std::memcpy(den_out, den.data(), x * z * sizeof(double));
    for (size_t i = 0; i < x; ++i)
        for (size_t k = 0; k < z; ++k) nu[i * z + k] *= nom[i * z + k] / (den[i * z + k] + eps);
    for (size_t i = 0; i < x; ++i) {
        double rs = 0.0;
        for (size_t k = 0; k < z; ++k) rs += nu[i * z + k];
        rs += eps;
        for (size_t k = 0; k < z; ++k) nu[i * z + k] /= rs;
    }
}

// update_mu (bld_appr.hpp:166-226): as update_nu with the roles of i and j
// exchanged, then the column normalisation mu_:j /= (sum_k mu_kj + eps).
void appr_update_mu(const Dims& d, double eps, const Appr& a, const double* nu, double* mu,
                    double* nom_out = nullptr, double* den_out = nullptr) {
    const size_t x = d.x, y = d.y, z = d.z;
    std::vector<double> nom(z * y, 0.0), den(z * y, 0.0);
    for (size_t j = 0; j < y; ++j)
        for (size_t k = 0; k < z; ++k)
            for (size_t i = 0; i < x; ++i)
                nom[k * y + j] += a.oa[i * y + j] * nu[i * z + k] * a.gp[d.at(i, j, k)];
    for (size_t j = 0; j < y; ++j)
        for (size_t k = 0; k < z; ++k)
            for (size_t i = 0; i < x; ++i)
                for (size_t c = 0; c < z; ++c)
                    nom[k * y + j] += a.oa2[i * y + j] * nu[i * z + k] * nu[i * z + c] * mu[c * y + j] *
                                      a.gm[i * z + c];
    for (size_t j = 0; j < y; ++j)
        for (size_t k = 0; k < z; ++k)
            for (size_t i = 0; i < x; ++i)
                den[k * y + j] += a.oa[i * y + j] * nu[i * z + k] * a.gm[i * z + k];
    for (size_t j = 0; j < y; ++j)
        for (size_t k = 0; k < z; ++k)
            for (size_t i = 0; i < x; ++i)
                for (size_t c = 0; c < z; ++c)
                    den[k * y + j] += a.oa2[i * y + j] * nu[i * z + k] * nu[i * z + c] * mu[c * y + j] *
                                      a.gp[d.at(i, j, c)];
    if (nom_out) std::memcpy(nom_out, nom.data(), z * y * sizeof(double));
    if (den_out) std::memcpy(den_out, den.data(), z * y * sizeof(double));
    for (size_t k = 0; k < z; ++k)
        for (size_t j = 0; j < y; ++j) mu[k * y + j] *= nom[k * y + j] / (den[k * y + j] + eps);
    for (size_t j = 0; j < y; ++j) {
        double cs = 0.0;
        for (size_t k = 0; k < z; ++k) cs += mu[k * y + j];
        cs += eps;
        for (size_t k = 0; k < z; ++k) mu[k * y + j] /= cs;
    }
}

void make_work(const Dims& d, Work& w, bool keep_grad_plus) {
    w.S_pjk.assign(d.y * d.z, 0.0);
    w.S_ipk.assign(d.x * d.z, 0.0);
    w.S_ijp.assign(d.x * d.y, 0.0);
    w.alpha_eph.assign(d.x * d.z, 0.0);
    w.beta_eph.assign(d.y * d.z, 0.0);
    w.psi_beta_eph.assign(d.y * d.z, 0.0);
    w.grad_minus.assign(d.x * d.z, 0.0);
    w.nom.assign(d.x * d.y, 0.0);
    w.denom.assign(d.x * d.y, 0.0);
    w.X_recip.assign(d.x * d.y, 0.0);
    if (keep_grad_plus) w.grad_plus.assign(d.x * d.y * d.z, 0.0);
}

// check_bld_params (util_details.hpp:20-28): X >= 0 everywhere. The alpha and
// beta lengths are fixed by the caller's (x, z) here.
int check_X(const double* X, size_t n) {
    for (size_t t = 0; t < n; ++t)
        if (!(X[t] >= 0.0)) return 1;
    return 0;
}

}  // namespace

extern "C" {

double oracle_psi(double x) { return psi_exact(x); }
double oracle_psi_appr(double x) { return psi_appr(x); }
double oracle_lngamma(double x) { return std::lgamma(x); }

void oracle_psi_vec(const double* x, double* out, size_t n, int appr) {
    for (size_t t = 0; t < n; ++t) out[t] = psi_sel(x[t], appr);
}

void oracle_tensor_sums(const double* S, size_t x, size_t y, size_t z,
                        double* S_pjk, double* S_ipk, double* S_ijp) {
    tensor_sums(S, Dims{x, y, z}, S_pjk, S_ipk, S_ijp);
}

// bld::bld_mult with an explicit initial S (S0). S0 == nullptr is rejected:
// the reference's init_S (bld_mult_funcs.hpp:36-74) draws from a GSL stream
// that depends on the thread count, so the oracle always takes S0 from the
// caller. Returns 0 on success, 1 for a parameter error.
int oracle_bld_mult(const double* X, size_t x, size_t y, size_t z,
                    const double* alpha, const double* beta, size_t max_iter,
                    int use_psi_appr, double eps, const double* S0, double* S,
                    int nthreads) {
    if (!X || !alpha || !beta || !S0 || !S) return 1;
    if (check_X(X, x * y)) return 1;
    set_threads(nthreads);
    const Dims d{x, y, z};
    Work w;
    make_work(d, w, false);
    // X_reciprocal (funcs:87-100)
    for (size_t t = 0; t < x * y; ++t) w.X_recip[t] = 1.0 / (X[t] + eps);
    std::memcpy(S, S0, x * y * z * sizeof(double));
    for (size_t eph = 0; eph < max_iter; ++eph)
        iterate(X, d, alpha, beta, use_psi_appr, eps, S, w, false);
    return 0;
}

// bld::bld_add (src/bld/bld_add.hpp:69-184) from an explicit S0 (the
// reference draws it inline from GSL's taus stream, :76-91, as init_S does).
// Returns 0 on success, 1 for a parameter error.
int oracle_bld_add(const double* X, size_t x, size_t y, size_t z, const double* alpha,
                   const double* beta, size_t max_iter, double eps, const double* S0, double* S,
                   int nthreads) {
    if (!X || !alpha || !beta || !S0 || !S) return 1;
    if (check_X(X, x * y)) return 1;
    set_threads(nthreads);
    const Dims d{x, y, z};
    Work w;
    make_work(d, w, false);
    std::memcpy(S, S0, x * y * z * sizeof(double));
    for (size_t eph = 0; eph < max_iter; ++eph) iterate_add(X, d, alpha, beta, eps, eph, S, w, nullptr);
    return 0;
}

// One bld_add iteration at iteration number eph; grad_S may be null.
int oracle_bld_add_step(const double* X, size_t x, size_t y, size_t z, const double* alpha,
                        const double* beta, double eps, size_t eph, const double* S_in,
                        double* S_out, double* grad_S) {
    if (!X || !alpha || !beta || !S_in || !S_out) return 1;
    if (check_X(X, x * y)) return 1;
    const Dims d{x, y, z};
    Work w;
    make_work(d, w, false);
    std::memcpy(S_out, S_in, x * y * z * sizeof(double));
    iterate_add(X, d, alpha, beta, eps, eph, S_out, w, grad_S);
    return 0;
}

// bld::bld_appr (src/bld/bld_appr.hpp:287-372) from explicit nu0 (x*z) and
// mu0 (z*y) (the reference draws nu's rows and mu's columns from
// Dirichlet(1) with GSL, :302-318). Writes nu, mu and S = (X/(nu mu + eps))
// nu_ik mu_kj (:366-369). Returns 0 on success, 1 for a parameter error.
int oracle_bld_appr(const double* X, size_t x, size_t y, size_t z, const double* alpha,
                    const double* beta, size_t max_iter, double eps, const double* nu0,
                    const double* mu0, double* S, double* nu, double* mu) {
    if (!X || !alpha || !beta || !nu0 || !mu0 || !S || !nu || !mu) return 1;
    if (check_X(X, x * y)) return 1;
    const Dims d{x, y, z};
    std::memcpy(nu, nu0, x * z * sizeof(double));
    std::memcpy(mu, mu0, z * y * sizeof(double));
    Appr a;
    for (size_t eph = 0; eph < max_iter; ++eph) {
        appr_helpers(X, d, alpha, beta, eps, nu, mu, a);
        appr_update_nu(d, eps, a, mu, nu);
        appr_helpers(X, d, alpha, beta, eps, nu, mu, a);
        appr_update_mu(d, eps, a, nu, mu);
    }
    for (size_t i = 0; i < x; ++i)
        for (size_t j = 0; j < y; ++j) {
            double h = 0.0;
            for (size_t k = 0; k < z; ++k) h += nu[i * z + k] * mu[k * y + j];
            const double oa = X[i * y + j] / (h + eps);
            for (size_t k = 0; k < z; ++k) S[d.at(i, j, k)] = oa * nu[i * z + k] * mu[k * y + j];
        }
    return 0;
}

// One half-step of bld_appr from (nu, mu): half 0 updates nu (update_nu),
// half 1 updates mu (update_mu). Writes the updated factor and that
// update's numerator and denominator (x*z for nu, z*y for mu).
int oracle_bld_appr_half(const double* X, size_t x, size_t y, size_t z, const double* alpha,
                         const double* beta, double eps, int half, const double* nu_in,
                         const double* mu_in, double* out, double* nom, double* den) {
    if (!X || !alpha || !beta || !nu_in || !mu_in || !out) return 1;
    if (check_X(X, x * y)) return 1;
    const Dims d{x, y, z};
    Appr a;
    appr_helpers(X, d, alpha, beta, eps, nu_in, mu_in, a);
    if (half == 0) {
        std::memcpy(out, nu_in, x * z * sizeof(double));
        appr_update_nu(d, eps, a, mu_in, out, nom, den);
    } else {
        std::memcpy(out, mu_in, z * y * sizeof(double));
        appr_update_mu(d, eps, a, nu_in, out, nom, den);
    }
    return 0;
}

// One iteration with every intermediate exposed (per-step parity checks).
// Any output pointer may be null.
int oracle_bld_mult_step(const double* X, size_t x, size_t y, size_t z,
                         const double* alpha, const double* beta,
                         int use_psi_appr, double eps, const double* S_in,
                         double* S_out, double* S_pjk, double* S_ipk,
                         double* S_ijp, double* grad_plus, double* grad_minus,
                         double* nom, double* denom) {
    if (!X || !alpha || !beta || !S_in || !S_out) return 1;
    if (check_X(X, x * y)) return 1;
    const Dims d{x, y, z};
    Work w;
    make_work(d, w, true);
    for (size_t t = 0; t < x * y; ++t) w.X_recip[t] = 1.0 / (X[t] + eps);
    std::memcpy(S_out, S_in, x * y * z * sizeof(double));
    iterate(X, d, alpha, beta, use_psi_appr, eps, S_out, w, true);
    auto put = [](double* dst, const std::vector<double>& v) {
        if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(double));
    };
    put(S_pjk, w.S_pjk);
    put(S_ipk, w.S_ipk);
    put(S_ijp, w.S_ijp);
    put(grad_plus, w.grad_plus);
    put(grad_minus, w.grad_minus);
    put(nom, w.nom);
    put(denom, w.denom);
    return 0;
}

// alloc_model::log_marginal_S (alloc_model_funcs.hpp:574-586) with its four
// terms. The loops keep the reference's accumulation order per term.
double oracle_log_marginal_S(const double* S, size_t xs, size_t ys, size_t zs,
                             const double* alpha, const double* beta, double a,
                             double b) {
    const long x = (long)xs, y = (long)ys, z = (long)zs;
    const Dims d{xs, ys, zs};
    // first term (:34-77)
    double alpha_sum = 0.0, lg_alpha_sum = 0.0;
    for (long i = 0; i < x; ++i) alpha_sum += alpha[i];
    for (long i = 0; i < x; ++i) lg_alpha_sum += std::lgamma(alpha[i]);
    double first = 0.0;
    for (long k = 0; k < z; ++k) {
        double sum = 0.0;
        for (long i = 0; i < x; ++i) {
            sum += alpha[i];
            for (long j = 0; j < y; ++j) sum += S[d.at(i, j, k)];
        }
        first -= std::lgamma(sum);
        for (long i = 0; i < x; ++i) {
            sum = alpha[i];
            for (long j = 0; j < y; ++j) sum += S[d.at(i, j, k)];
            first += std::lgamma(sum);
        }
    }
    first += std::lgamma(alpha_sum) * z - lg_alpha_sum * z;
    // second term (:91-132)
    double beta_sum = 0.0, lg_beta_sum = 0.0;
    for (long k = 0; k < z; ++k) beta_sum += beta[k];
    for (long k = 0; k < z; ++k) lg_beta_sum += std::lgamma(beta[k]);
    double second = 0.0;
    for (long j = 0; j < y; ++j) {
        double sum = 0.0;
        for (long k = 0; k < z; ++k) {
            sum += beta[k];
            for (long i = 0; i < x; ++i) sum += S[d.at(i, j, k)];
        }
        second -= std::lgamma(sum);
        for (long k = 0; k < z; ++k) {
            sum = beta[k];
            for (long i = 0; i < x; ++i) sum += S[d.at(i, j, k)];
            second += std::lgamma(sum);
        }
    }
    second += std::lgamma(beta_sum) * y - lg_beta_sum * y;
    // third term (:146-169)
    double third = 0.0;
    for (long j = 0; j < y; ++j) {
        double sum = a;
        for (long i = 0; i < x; ++i)
            for (long k = 0; k < z; ++k) sum += S[d.at(i, j, k)];
        third += std::lgamma(sum);
        third -= sum * std::log(b + 1);
    }
    third += -std::lgamma(a) * y + a * std::log(b) * y;
    // fourth term (:179-194)
    double fourth = 0.0;
    for (size_t t = 0; t < xs * ys * zs; ++t) fourth += std::lgamma(S[t] + 1);
    return first + second + third - fourth;
}

// util::sparseness (util.hpp:270-295).
double oracle_sparseness(const double* S, size_t x, size_t y, size_t z) {
    double sum = 0.0, sq = 0.0;
    for (size_t t = 0; t < x * y * z; ++t) {
        sum += S[t];
        sq += S[t] * S[t];
    }
    if (sq < std::numeric_limits<double>::epsilon())
        return std::numeric_limits<double>::max();
    const double frob = std::sqrt(sq);
    const double axis_mult = std::sqrt((double)(x * y * z));
    return (axis_mult - sum / frob) / (axis_mult - 1);
}

// bld::bld_fact (bld_fact.hpp:36-80): W (x*z), H (z*y), L (y).
void oracle_bld_fact(const double* S, size_t x, size_t y, size_t z,
                     const double* alpha, const double* beta, double a,
                     double b, double eps, double* W, double* H, double* L) {
    const Dims d{x, y, z};
    std::vector<double> S_pjk(y * z), S_ipk(x * z);
    tensor_sums(S, d, S_pjk.data(), S_ipk.data(), nullptr);
    std::vector<double> S_pjp(y, 0.0);
    for (size_t j = 0; j < y; ++j)
        for (size_t k = 0; k < z; ++k) S_pjp[j] += S_pjk[j * z + k];
    for (size_t i = 0; i < x; ++i)
        for (size_t k = 0; k < z; ++k) W[i * z + k] = alpha[i] + S_ipk[i * z + k] - 1;
    for (size_t k = 0; k < z; ++k)
        for (size_t j = 0; j < y; ++j) H[k * y + j] = beta[k] + S_pjk[j * z + k] - 1;
    for (size_t j = 0; j < y; ++j) L[j] = (a + S_pjp[j] - 1) / (b + 1 + eps);
    for (size_t k = 0; k < z; ++k) {
        double cs = 0.0;
        for (size_t i = 0; i < x; ++i) cs += W[i * z + k];
        cs += eps;
        for (size_t i = 0; i < x; ++i) W[i * z + k] /= cs;
    }
    for (size_t j = 0; j < y; ++j) {
        double cs = 0.0;
        for (size_t k = 0; k < z; ++k) cs += H[k * y + j];
        cs += eps;
        for (size_t k = 0; k < z; ++k) H[k * y + j] /= cs;
    }
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
