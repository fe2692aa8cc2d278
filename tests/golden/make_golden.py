"""Writes tests/golden/golden.json: high-precision reference vectors for the
BLD solver's numerics, computed with mpmath at 50 significant digits.

Nothing here runs the reference code (it cannot be built on this machine:
Eigen3 and GSL are absent). The vectors come from:
  * mpmath.digamma / mpmath.loggamma (the exact functions gsl_sf_psi and
    gsl_sf_lngamma evaluate);
  * util::psi_appr's closed form (src/util/util.hpp:440-514: recurrence
    shift below 6, then the x^-14 series) evaluated in mpmath;
  * one and five multiplicative iterations (bld_mult.hpp:123-147 and
    bld_mult_funcs.hpp:130-327) on tiny tensors, in mpmath arithmetic;
  * log_marginal_S (alloc_model_funcs.hpp:34-194, 574-586) in mpmath;
  * the 8x10 fixture of test/bld/bld_mult/test_bld_mult.cpp:50-58 with a
    committed initial tensor S0 (the reference draws S0 from a GSL stream
    whose values depend on the thread count, so S0 is fixed here instead).

Run: python tests/golden/make_golden.py   (takes a few seconds)
"""
import json
import os

import mpmath
import numpy as np

mpmath.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))


def f(v):
    return float(v)


def psi_appr_mp(x):
    x = mpmath.mpf(x)
    if x < 1:
        n = 6
    elif x < 6:
        n = 6 - int(mpmath.floor(x))
    else:
        n = 0
    extra = sum((1 / (x + m) for m in range(n)), mpmath.mpf(0))
    t = x + n
    series = (-360360 * t**13 - 60060 * t**12 + 6006 * t**10 - 2860 * t**8 + 3003 * t**6
              - 5460 * t**4 + 15202 * t**2 - 60060) / (720720 * t**14)
    return mpmath.log(t) + series - extra


def psi_points():
    fixed = [1e-8, 1e-5, 1e-3, 0.01, 0.05, 0.1, 0.25, 0.5, 0.75, 1.0, 1.25,
             1.4616321449683623, 1.5, 2.0, 2.5, 3.0, 4.0, 5.0, 5.999, 6.0, 7.5, 9.99, 10.0,
             10.01, 12.5, 20.0, 33.3, 50.0, 100.0, 1e3, 1e4, 1e6, 1e10]
    rng = np.random.default_rng(20240611)
    rand = list(10 ** rng.uniform(-6, 6, 200)) + list(rng.uniform(0, 12, 200))
    return fixed + [float(v) for v in rand]


def iterate_mp(X, S, alpha, beta, psi, eps=mpmath.mpf("1e-50")):
    """One bld_mult iteration (bld_mult.hpp:123-147) in mpmath; returns the
    new S and the intermediates."""
    x, y, z = len(S), len(S[0]), len(S[0][0])
    S_pjk = [[sum(S[i][j][k] for i in range(x)) for k in range(z)] for j in range(y)]
    S_ipk = [[sum(S[i][j][k] for j in range(y)) for k in range(z)] for i in range(x)]
    S_ijp = [[sum(S[i][j][k] for k in range(z)) for j in range(y)] for i in range(x)]
    alpha_eph = [[S_ipk[i][k] + alpha[i] for k in range(z)] for i in range(x)]
    beta_eph = [[S_pjk[j][k] + beta[k] for k in range(z)] for j in range(y)]
    gp = [[[psi(beta_eph[j][k]) - psi(S[i][j][k] + 1) for k in range(z)] for j in range(y)]
          for i in range(x)]
    colsum = [sum(alpha_eph[i][k] for i in range(x)) for k in range(z)]
    gm = [[psi(colsum[k]) - psi(alpha_eph[i][k]) for k in range(z)] for i in range(x)]
    Xr = [[1 / (X[i][j] + eps) for j in range(y)] for i in range(x)]
    nom = [[sum(S[i][j][k] * Xr[i][j] * gm[i][k] for k in range(z)) for j in range(y)]
           for i in range(x)]
    den = [[sum(S[i][j][k] * Xr[i][j] * gp[i][j][k] for k in range(z)) for j in range(y)]
           for i in range(x)]
    Snew = [[[S[i][j][k] * (gp[i][j][k] + nom[i][j]) * (X[i][j] / (S_ijp[i][j] + eps))
              / (gm[i][k] + den[i][j] + eps) for k in range(z)] for j in range(y)]
            for i in range(x)]
    inter = {"S_pjk": S_pjk, "S_ipk": S_ipk, "S_ijp": S_ijp, "grad_plus": gp,
             "grad_minus": gm, "nom": nom, "denom": den}
    return Snew, inter


def log_marginal_mp(S, alpha, beta, a, b):
    """alloc_model::log_marginal_S (alloc_model_funcs.hpp:574-586)."""
    lg = mpmath.loggamma
    x, y, z = len(S), len(S[0]), len(S[0][0])
    a, b = mpmath.mpf(a), mpmath.mpf(b)
    sa, sb = sum(alpha), sum(beta)
    first = z * (lg(sa) - sum(lg(v) for v in alpha))
    for k in range(z):
        first -= lg(sa + sum(S[i][j][k] for i in range(x) for j in range(y)))
        for i in range(x):
            first += lg(alpha[i] + sum(S[i][j][k] for j in range(y)))
    second = y * (lg(sb) - sum(lg(v) for v in beta))
    for j in range(y):
        second -= lg(sb + sum(S[i][j][k] for i in range(x) for k in range(z)))
        for k in range(z):
            second += lg(beta[k] + sum(S[i][j][k] for i in range(x)))
    third = y * (-lg(a) + a * mpmath.log(b))
    for j in range(y):
        s = a + sum(S[i][j][k] for i in range(x) for k in range(z))
        third += lg(s) - s * mpmath.log(b + 1)
    fourth = sum(lg(S[i][j][k] + 1) for i in range(x) for j in range(y) for k in range(z))
    return first + second + third - fourth


def tolist3(S):
    return [[[f(v) for v in row] for row in mat] for mat in S]


def tolist2(M):
    return [[f(v) for v in row] for row in M]


def tiny_case(name, X, alpha, beta, seed, a, b):
    rng = np.random.default_rng(seed)
    x, y = len(X), len(X[0])
    z = len(beta)
    S0 = np.zeros((x, y, z))
    for i in range(x):
        for j in range(y):
            S0[i, j] = X[i][j] * rng.dirichlet(np.ones(z))
    out = {"name": name, "X": X, "alpha": alpha, "beta": beta, "a": a, "b": b,
           "S0": S0.tolist()}
    for mode, psi in (("exact", mpmath.digamma), ("appr", psi_appr_mp)):
        Xm = [[mpmath.mpf(v) for v in row] for row in X]
        am = [mpmath.mpf(v) for v in alpha]
        bm = [mpmath.mpf(v) for v in beta]
        S = [[[mpmath.mpf(float(v)) for v in S0[i, j]] for j in range(y)] for i in range(x)]
        S1, inter = iterate_mp(Xm, S, am, bm, psi)
        S5 = S1
        for _ in range(4):
            S5, _ = iterate_mp(Xm, S5, am, bm, psi)
        out[mode] = {
            "S1": tolist3(S1), "S5": tolist3(S5),
            "S_pjk": tolist2(inter["S_pjk"]), "S_ipk": tolist2(inter["S_ipk"]),
            "S_ijp": tolist2(inter["S_ijp"]), "grad_plus": tolist3(inter["grad_plus"]),
            "grad_minus": tolist2(inter["grad_minus"]), "nom": tolist2(inter["nom"]),
            "denom": tolist2(inter["denom"]),
        }
    S0m = [[[mpmath.mpf(float(v)) for v in S0[i, j]] for j in range(y)] for i in range(x)]
    out["log_marginal_S0"] = f(log_marginal_mp(S0m, [mpmath.mpf(v) for v in alpha],
                                               [mpmath.mpf(v) for v in beta], a, b))
    S5m = [[[mpmath.mpf(v) for v in row] for row in mat] for mat in out["exact"]["S5"]]
    out["log_marginal_S5"] = f(log_marginal_mp(S5m, [mpmath.mpf(v) for v in alpha],
                                               [mpmath.mpf(v) for v in beta], a, b))
    return out


def main():
    pts = psi_points()
    golden = {
        "psi": {"x": pts, "psi": [f(mpmath.digamma(mpmath.mpf(v))) for v in pts]},
        "psi_neg": {"x": [-0.5, -1.5, -2.25, -7.75],
                    "psi": [f(mpmath.digamma(mpmath.mpf(v))) for v in (-0.5, -1.5, -2.25, -7.75)]},
        "psi_appr": {"x": pts, "psi_appr": [f(psi_appr_mp(v)) for v in pts]},
        "lngamma": {"x": [v for v in pts if v < 1e8],
                    "lngamma": [f(mpmath.loggamma(mpmath.mpf(v))) for v in pts if v < 1e8]},
    }
    cases = []
    cases.append(tiny_case("3x4x2", [[2.0, 0.0, 5.0, 1.0], [7.0, 3.0, 1.0, 4.0],
                                     [1.0, 1.0, 0.0, 9.0]],
                           [0.5, 1.0, 2.0], [1.0, 3.0], 7, 40.0, 1.0))
    cases.append(tiny_case("5x3x3", [[4.0, 0.0, 2.0], [0.0, 0.0, 0.0], [11.0, 6.0, 1.0],
                                     [3.0, 0.0, 8.0], [1.0, 2.0, 0.0]],
                           [1.0, 0.05, 2.5, 1.0, 0.7], [1.0, 3.0, 0.7], 11, 10.0, 2.0))
    golden["iterations"] = cases
    # 8x10 fixture (test/bld/bld_mult/test_bld_mult.cpp:50-58)
    X8 = [7., 11., 5., 8., 8., 8., 7., 6., 3., 9., 3., 7., 3., 1., 0., 6.,
          1., 0., 1., 4., 1., 11., 4., 5., 3., 8., 5., 3., 3., 3., 0., 20.,
          11., 16., 14., 26., 16., 23., 16., 15., 0., 1., 0., 2., 2., 2., 0.,
          1., 0., 1., 3., 4., 1., 0., 4., 0., 2., 0., 1., 5., 4., 4., 2., 2.,
          3., 4., 0., 2., 0., 5., 4., 3., 3., 1., 2., 6., 5., 7., 4., 2.]
    X8 = np.array(X8).reshape(8, 10)
    rng = np.random.default_rng(12)
    S0_8 = X8[:, :, None] * rng.dirichlet(np.ones(3), size=(8, 10))
    golden["fixture_8x10"] = {
        "X": X8.tolist(), "z": 3, "a": 40.0, "b": 1.0, "alpha": [1.0] * 8, "beta": [1.0] * 3,
        "S0": S0_8.tolist(), "max_iter": 2000,
        "thresholds": {"log_marginal_min": -262.0, "sparseness_min": 0.64,
                       "isapprox_exact": 1e-10, "isapprox_appr": 1e-4},
        "source": "test/bld/bld_mult/test_bld_mult.cpp:48-85",
    }
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump(golden, fh)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
