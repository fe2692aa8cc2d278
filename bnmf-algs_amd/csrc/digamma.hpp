// digamma.hpp — digamma (psi) functions evaluated in registers on the GPU.
//
// Four forms are provided:
//  * psi_f64        : the exact digamma in double, the role GSL's gsl_sf_psi
//                     plays for the reference CPU solver
//                     (src/bld/util_details.hpp:71-73, bld_mult.hpp:119-120).
//                     On [1, 10) (almost every argument of a solve) a
//                     rational fit on [1, 2) reached by downward recurrence
//                     (one division, no log); elsewhere the recurrence
//                     psi(x) = psi(x+1) - 1/x (Abramowitz & Stegun 6.3.5)
//                     lifts x to >= 10, the shift sum accumulated as one
//                     fraction so that a single division is paid, then the
//                     asymptotic series ln x - 1/(2x) - sum B_2n/(2n x^2n)
//                     (A&S 6.3.18) is summed to the x^-16 term.
//  * psi_appr_ref   : util::psi_appr (src/util/util.hpp:440-514), the
//                     reference's approximate digamma: the same case split
//                     and the same closed-form series, evaluated in T. The
//                     steps have to match the reference's for the FP64
//                     approximate mode to agree with it, so they are kept in
//                     the reference's order.
//  * psi_ge1_f32 / psi_f32 : FP32 digamma for the FP32-bulk solver.
//                     psi(x) = ln x + Q(1/x) where Q(u) = psi(1/u) + ln u is a
//                     degree-8 polynomial fitted (weighted least squares on
//                     Chebyshev nodes, max |error| 3.3e-9 on u in [0,1]) to
//                     mpmath's digamma. One reciprocal, one log and eight
//                     FMAs; no branches for x >= 1. Accuracy is limited by the
//                     float32 log (about 1.5e-7 relative), i.e. the same as
//                     any float32 evaluation of psi_appr.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace bnmf {

// ---------------------------------------------------------------- FP64 exact
// psi(1 + t) for t in [0, 1): P(t) / Q(t) with P, Q of degree 6, fitted for
// this project to mpmath's digamma (Loeb's linearised least squares with
// Lawson reweighting towards the minimax error; max |error| 2.3e-19 in exact
// arithmetic, about one ulp of the result when evaluated in double).
__host__ __device__ inline double psi1p_rat(double t, double* q_out = nullptr) {
    double p = 2.2740920435799464e-04;
    p = fma(p, t, 1.2358749079795044e-02);
    p = fma(p, t, 1.6310050654079108e-01);
    p = fma(p, t, 7.8211906069224120e-01);
    p = fma(p, t, 1.3871517644465865e+00);
    p = fma(p, t, 4.2674224942732580e-01);
    p = fma(p, t, -5.7721566490153290e-01);
    double q = 4.0209462601665044e-05;
    q = fma(q, t, 3.4653371266403213e-03);
    q = fma(q, t, 6.6582844787605800e-02);
    q = fma(q, t, 4.8134849233954260e-01);
    q = fma(q, t, 1.5286529550625810e+00);
    q = fma(q, t, 2.1104621573787530e+00);
    q = fma(q, t, 1.0);
    if (q_out) {
        *q_out = q;
        return p;
    }
    return p / q;
}

// Any argument: the recurrence to x >= 10 and the asymptotic series, the
// reflection formula for x <= 0.
__host__ __device__ inline double psi_f64_general(double x) {
    if (!(x > 0.0)) {
        if (x != x) return x;
        if (x == floor(x)) return __builtin_nan("");
        // reflection psi(1-x) - psi(x) = pi cot(pi x)   (A&S 6.3.7)
        const double pi = 3.14159265358979323846264338327950288;
        const double xr = 1.0 - x;
        // 1 - x > 0 here; the shift below handles it.
        double num = 0.0, den = 1.0, t = xr;
        while (t < 10.0) {
            num = num * t + den;
            den = den * t;
            t += 1.0;
        }
        const double r = 1.0 / t, r2 = r * r;
        const double series =
            r2 * (1.0 / 12 - r2 * (1.0 / 120 - r2 * (1.0 / 252 - r2 * (1.0 / 240 - r2 * (1.0 / 132 - r2 * (691.0 / 32760 - r2 * (1.0 / 12 - r2 * (3617.0 / 8160))))))));
        const double psi_1mx = (log(t) - 0.5 * r - series) - num / den;
        return psi_1mx - pi / tan(pi * x);
    }
    // sum_{m<n} 1/(x+m) accumulated as num/den: each step adds 1/t to num/den
    double num = 0.0, den = 1.0;
    while (x < 10.0) {
        num = num * x + den;
        den = den * x;
        x += 1.0;
    }
    const double r = 1.0 / x;
    const double r2 = r * r;
    const double series =
        r2 * (1.0 / 12 -
              r2 * (1.0 / 120 -
                    r2 * (1.0 / 252 -
                          r2 * (1.0 / 240 -
                                r2 * (1.0 / 132 -
                                      r2 * (691.0 / 32760 -
                                            r2 * (1.0 / 12 - r2 * (3617.0 / 8160))))))));
    return (log(x) - 0.5 * r - series) - num / den;
}

#if defined(__HIP_DEVICE_COMPILE__)
// n / d for d of moderate magnitude (|d| in [2^-1000, 2^1000]): the hardware
// reciprocal refined by two Newton steps, then one correction of the
// quotient from its residual. Within one ulp of n / d; an exact-arithmetic
// emulation of the steps gave the correctly rounded quotient for all of
// 20,000 random cases. 1.6 % faster than the IEEE division sequence in the
// FP64 step pass, where both divisions per element use it.
__device__ inline double div_moderate(double n, double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    const double q = n * r;
    return fma(fma(-d, q, n), r, q);
}
#else
__host__ __device__ inline double div_moderate(double n, double d) { return n / d; }
#endif

// psi(x) for 1 <= x < 10 (the caller guarantees the range): the recurrence
// psi(x) = psi(x - n) + sum_{m=1..n} 1/(x - m) (Abramowitz & Stegun 6.3.5)
// brings x down to [1, 2), where psi1p_rat applies; the shift sum is kept as
// one fraction num/den and joined to P/Q, so one division is paid. The
// shifted argument y = x - m steps down by 1 (exactly: y >= 1) while y >= 1,
// so the last y is t = x - floor(x), psi1p_rat's argument.
__host__ __device__ inline double psi_1to10(double x) {
    double y = x - 1.0;
    double num = 0.0, den = 1.0;
    while (y >= 1.0) {
        num = fma(num, y, den);
        den = den * y;
        y -= 1.0;
    }
    double q;
    const double p = psi1p_rat(y, &q);
    return div_moderate(fma(p, den, q * num), q * den);
}

// The exact digamma of the FP64 forms: psi_1to10 on [1, 10), psi_f64_general
// elsewhere. Every FP64 digamma (prep's psi(beta_eph), psi(alpha_eph) and
// psi(A), the pass's psi(S + 1)) is this one function, so equal arguments
// give equal values, as in the reference (a single row with beta = 1 has
// psi(beta_eph) = psi(S + 1) exactly, hence a zero gradient).
__host__ __device__ inline double psi_f64(double x) {
    return (x >= 1.0 && x < 10.0) ? psi_1to10(x) : psi_f64_general(x);
}

// psi(s + 1), the per-element term of the FP64 update, at the rounded
// argument x = s + 1 as the reference evaluates it.
__host__ __device__ inline double psi_sp1_f64(double s) { return psi_f64(s + 1.0); }

// ------------------------------------------------- reference psi_appr form
template <typename Real>
__host__ __device__ inline Real psi_appr_ref(Real x) {
    Real extra = 0;
    if (x < 1) {
        const Real a = x + 1, b = x + 2, c = x + 3, d = x + 4, e = x + 5;
        const Real ab = a * b, cd = c * d, ex = e * x;
        extra = ((a + b) * cd * ex + ab * d * ex + ab * c * ex + ab * cd * x + ab * cd * e) /
                (ab * cd * ex);
        x += 6;
    } else if (x < 2) {
        const Real a = x + 1, b = x + 2, c = x + 3, d = x + 4;
        const Real ab = a * b, cd = c * d, dx = d * x;
        extra = ((a + b) * c * dx + ab * dx + ab * c * x + ab * cd) / (ab * cd * x);
        x += 5;
    } else if (x < 3) {
        const Real a = x + 1, b = x + 2, c = x + 3;
        const Real ab = a * b, cx = c * x;
        extra = ((a + b) * cx + (c + x) * ab) / (ab * cx);
        x += 4;
    } else if (x < 4) {
        const Real a = x + 1, b = x + 2;
        const Real ab = a * b;
        extra = ((a + b) * x + ab) / (ab * x);
        x += 3;
    } else if (x < 5) {
        const Real a = x + 1;
        extra = (a + x) / (a * x);
        x += 2;
    } else if (x < 6) {
        extra = 1 / x;
        x += 1;
    }
    const Real x2 = x * x, x4 = x2 * x2, x6 = x4 * x2, x8 = x6 * x2;
    const Real x10 = x8 * x2, x12 = x10 * x2, x13 = x12 * x, x14 = x13 * x;
    const Real res = log(x) + (-360360 * x13 - 60060 * x12 + 6006 * x10 - 2860 * x8 +
                               3003 * x6 - 5460 * x4 + 15202 * x2 - 60060) /
                                  (720720 * x14);
    return res - extra;
}

// ------------------------------------------------------------------- FP32
// Q(u) = psi(1/u) + ln(u) = sum_{n=1..8} q_n u^n on u in [0, 1].
#define BNMF_PSI32_Q1 (-0.5000003576278687f)
#define BNMF_PSI32_Q2 (-0.08332289755344391f)
#define BNMF_PSI32_Q3 (-0.0001106799827539362f)
#define BNMF_PSI32_Q4 (0.008870677091181278f)
#define BNMF_PSI32_Q5 (-0.001238060649484396f)
#define BNMF_PSI32_Q6 (-0.0031849101651459932f)
#define BNMF_PSI32_Q7 (0.002303620334714651f)
#define BNMF_PSI32_Q8 (-0.0005330853164196014f)

__host__ __device__ inline float psi32_q(float u) {
    float p = BNMF_PSI32_Q8;
    p = fmaf(p, u, BNMF_PSI32_Q7);
    p = fmaf(p, u, BNMF_PSI32_Q6);
    p = fmaf(p, u, BNMF_PSI32_Q5);
    p = fmaf(p, u, BNMF_PSI32_Q4);
    p = fmaf(p, u, BNMF_PSI32_Q3);
    p = fmaf(p, u, BNMF_PSI32_Q2);
    p = fmaf(p, u, BNMF_PSI32_Q1);
    return p * u;
}

#if defined(__HIP_DEVICE_COMPILE__)
// x >= 1 (the hot-path argument S_ijk + 1): hardware reciprocal and log2
// (x is never denormal here, so no input scaling), ln 2 folded into the
// final FMA.
__device__ inline float psi_ge1_f32(float x) {
    const float u = __builtin_amdgcn_rcpf(x);
    return fmaf(__builtin_amdgcn_logf(x), 0.69314718055994531f, psi32_q(u));
}
#else
__host__ __device__ inline float psi_ge1_f32(float x) { return logf(x) + psi32_q(1.0f / x); }
#endif

// any x > 0 (prep arguments S_i+k + alpha_i, S_+jk + beta_k, their column sums)
__host__ __device__ inline float psi_f32(float x) {
    float shift = 0.0f;
    if (x < 1.0f) {  // psi(x) = psi(x + 1) - 1/x
        shift = 1.0f / x;
        x += 1.0f;
    }
    return psi_ge1_f32(x) - shift;
}

// Selected by the solver's numeric mode.
enum PsiMode : int { PSI_F32 = 0, PSI_F64_EXACT = 1, PSI_F64_APPR = 2 };

template <int MODE, typename T>
__host__ __device__ inline T psi_sel(T x);
template <>
__host__ __device__ inline float psi_sel<PSI_F32, float>(float x) { return psi_f32(x); }
template <>
__host__ __device__ inline double psi_sel<PSI_F64_EXACT, double>(double x) { return psi_f64(x); }
template <>
__host__ __device__ inline double psi_sel<PSI_F64_APPR, double>(double x) { return psi_appr_ref<double>(x); }

// psi(s + 1) for s >= 0, the per-element hot term.
template <int MODE, typename T>
__host__ __device__ inline T psi_sp1(T s);
template <>
__host__ __device__ inline float psi_sp1<PSI_F32, float>(float s) {
    // s + 1 >= 1 whenever s >= 0; a negative s (only reachable with beta < 1)
    // takes the general path.
    const float a = s + 1.0f;
    return a >= 1.0f ? psi_ge1_f32(a) : psi_f32(a);
}
template <>
__host__ __device__ inline double psi_sp1<PSI_F64_EXACT, double>(double s) { return psi_sp1_f64(s); }
template <>
__host__ __device__ inline double psi_sp1<PSI_F64_APPR, double>(double s) { return psi_appr_ref<double>(s + 1); }

}  // namespace bnmf
