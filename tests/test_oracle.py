"""Pins the FP64 CPU oracle (oracle/bld_oracle.cpp) before it is trusted.

Each oracle function is checked against something independent of its own
formula: mpmath golden vectors (tests/golden/golden.json), closed forms of
the digamma, brute-force sums, the gradient identity linking the update's
grad_plus/grad_minus to log_marginal_S, fixed points of the update, and the
reference's own known-answer tests (test/bld/bld_mult/test_bld_mult.cpp,
test/util/test_util.cpp, test/allocation_model/test_alloc_model_funcs.cpp).
"""
import math

import numpy as np
import pytest

EULER = 0.57721566490153286060651209008240243


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# ------------------------------------------------------------------ digamma
def test_psi_exact_matches_mpmath(oracle, golden):
    xs = np.array(golden["psi"]["x"])
    ref = np.array(golden["psi"]["psi"])
    got = oracle.psi_vec(xs)
    err = np.abs(got - ref)
    # the shift sum cancels against the series near the root x0 = 1.4616
    tol = 4e-15 * np.maximum(1.0, np.abs(ref))
    assert np.all(err <= tol), (xs[np.argmax(err / tol)], err.max())


def test_psi_exact_closed_forms(oracle):
    assert oracle.psi(1.0) == pytest.approx(-EULER, abs=1e-15)
    assert oracle.psi(0.5) == pytest.approx(-EULER - 2 * math.log(2), abs=2e-15)
    h = 0.0
    for n in range(1, 40):
        h += 1.0 / n
        assert oracle.psi(n + 1.0) == pytest.approx(h - EULER, abs=2e-15 * max(1, h))
    rng = np.random.default_rng(3)
    for v in rng.uniform(0.01, 30, 200):  # psi(x+1) - psi(x) = 1/x
        assert oracle.psi(v + 1) - oracle.psi(v) == pytest.approx(1 / v, rel=1e-12, abs=1e-13)


def test_psi_exact_reflection(oracle, golden):
    for v, ref in zip(golden["psi_neg"]["x"], golden["psi_neg"]["psi"]):
        assert oracle.psi(v) == pytest.approx(ref, rel=1e-13, abs=1e-13)
    assert math.isnan(oracle.psi(0.0)) and math.isnan(oracle.psi(-3.0))


def test_psi_appr_matches_formula(oracle, golden):
    xs = np.array(golden["psi_appr"]["x"])
    ref = np.array(golden["psi_appr"]["psi_appr"])
    got = oracle.psi_vec(xs, appr=True)
    assert np.all(np.abs(got - ref) <= 2e-15 * np.maximum(1.0, np.abs(ref)))


def test_psi_appr_reference_bounds(oracle):
    # test/util/test_util.cpp:309-341: |psi_appr - psi| <= 1e-10 on (0, 1)
    # with step 1e-5 and <= 1e-12 on [1, 100) with step 1e-3
    xa = np.arange(1, 100000) * 1e-5
    assert np.abs(oracle.psi_vec(xa, True) - oracle.psi_vec(xa)).max() <= 1e-10
    xb = 1.0 + np.arange(0, 99000) * 1e-3
    assert np.abs(oracle.psi_vec(xb, True) - oracle.psi_vec(xb)).max() <= 1e-12


def test_lngamma_matches_mpmath(oracle, golden):
    for v, ref in zip(golden["lngamma"]["x"], golden["lngamma"]["lngamma"]):
        assert oracle.lib().oracle_lngamma(v) == pytest.approx(ref, rel=1e-14, abs=1e-14)


# ----------------------------------------------------------------- axis sums
def test_tensor_sums_bruteforce(oracle):
    rng = np.random.default_rng(0)
    S = rng.random((7, 9, 5)) * 10
    pjk, ipk, ijp = oracle.tensor_sums(S)
    for j in range(9):
        for k in range(5):
            assert pjk[j, k] == pytest.approx(math.fsum(S[:, j, k]), rel=1e-15)
    for i in range(7):
        for k in range(5):
            assert ipk[i, k] == pytest.approx(math.fsum(S[i, :, k]), rel=1e-15)
        for j in range(9):
            assert ijp[i, j] == pytest.approx(math.fsum(S[i, j, :]), rel=1e-15)
    tot = math.fsum(S.ravel())
    for part in (pjk, ipk, ijp):
        assert math.fsum(part.ravel()) == pytest.approx(tot, rel=1e-14)


# ----------------------------------------------------------------- iteration
@pytest.mark.parametrize("mode", ["exact", "appr"])
@pytest.mark.parametrize("case", [0, 1])
def test_iteration_matches_mpmath(oracle, golden, mode, case):
    c = golden["iterations"][case]
    X, S0 = np.array(c["X"]), np.array(c["S0"])
    appr = mode == "appr"
    step = oracle.bld_mult_step(X, c["alpha"], c["beta"], S0, use_psi_appr=appr)
    g = c[mode]
    for key in ("S_pjk", "S_ipk", "S_ijp", "grad_plus", "grad_minus", "nom", "denom"):
        assert _rel(step[key], g[key]) < 1e-13, key
    assert _rel(step["S"], g["S1"]) < 1e-13
    S5 = oracle.bld_mult(X, len(c["beta"]), c["alpha"], c["beta"], S0, max_iter=5,
                         use_psi_appr=appr)
    assert _rel(S5, g["S5"]) < 1e-12


def test_gradient_identity(oracle):
    """grad_plus - grad_minus is d log p(S) / dS_ijk up to a term c_j that
    depends on j only (the derivative of the L and S_+j+ terms), so central
    differences of log_marginal_S pin the index layout of both gradients."""
    rng = np.random.default_rng(5)
    x, y, z = 3, 4, 3
    alpha = np.array([0.5, 1.0, 2.0])
    beta = np.array([1.0, 3.0, 0.7])
    a, b = 7.0, 2.0
    S = rng.uniform(0.5, 4.0, (x, y, z))
    X = S.sum(2)
    st = oracle.bld_mult_step(X, alpha, beta, S)
    d = np.empty((x, y, z))
    h = 1e-5
    for i in range(x):
        for j in range(y):
            for k in range(z):
                Sp, Sm = S.copy(), S.copy()
                Sp[i, j, k] += h
                Sm[i, j, k] -= h
                d[i, j, k] = (oracle.log_marginal_S(Sp, alpha, beta, a, b) -
                              oracle.log_marginal_S(Sm, alpha, beta, a, b)) / (2 * h)
    resid = d - (st["grad_plus"] - st["grad_minus"][:, None, :])
    spj = S.sum(axis=(0, 2))
    cj = oracle.psi_vec(a + spj) - oracle.psi_vec(beta.sum() + spj) - math.log(b + 1)
    assert np.abs(resid - cj[None, :, None]).max() < 1e-7


def test_z1_is_fixed_point(oracle):
    rng = np.random.default_rng(1)
    X = rng.integers(0, 9, (6, 5)).astype(float)
    S = oracle.bld_mult(X, 1, np.ones(6), np.ones(1), X[:, :, None].copy(), max_iter=50)
    assert np.allclose(S[:, :, 0], X, rtol=1e-13, atol=0)


def test_zero_matrix_gives_zero(oracle):
    # test_bld_mult.cpp:36-46
    X = np.zeros((8, 10))
    S = oracle.bld_mult(X, 3, np.ones(8), np.ones(3), np.zeros((8, 10, 3)), max_iter=100)
    assert np.all(S == 0)


def test_zero_fibers_stay_zero(oracle):
    rng = np.random.default_rng(2)
    X = rng.integers(0, 4, (9, 7)).astype(float)
    X[2, :] = 0
    X[:, 3] = 0
    S0 = X[:, :, None] * rng.dirichlet(np.ones(4), size=(9, 7))
    S = oracle.bld_mult(X, 4, np.ones(9), np.ones(4), S0, max_iter=40)
    assert np.all(S[X == 0] == 0)
    assert np.all(np.isfinite(S))


def test_rejects_negative_X(oracle):
    X = np.full((4, 3), 5.0)
    X[0, 0] = -5
    with pytest.raises(ValueError):
        oracle.bld_mult(X, 2, np.ones(4), np.ones(2), np.ones((4, 3, 2)), max_iter=1)


@pytest.mark.parametrize("appr", [False, True])
def test_fixture_8x10_known_answer(oracle, golden, appr):
    # test_bld_mult.cpp:48-85 thresholds after 2000 iterations
    fx = golden["fixture_8x10"]
    X, S0 = np.array(fx["X"]), np.array(fx["S0"])
    S = oracle.bld_mult(X, fx["z"], fx["alpha"], fx["beta"], S0, max_iter=fx["max_iter"],
                        use_psi_appr=appr)
    th = fx["thresholds"]
    assert oracle.log_marginal_S(S, fx["alpha"], fx["beta"], fx["a"], fx["b"]) >= th["log_marginal_min"]
    assert oracle.sparseness(S) >= th["sparseness_min"]
    sij = S.sum(2)
    rel = np.linalg.norm(X - sij) / min(np.linalg.norm(X), np.linalg.norm(sij))
    assert rel <= (th["isapprox_appr"] if appr else th["isapprox_exact"])


# -------------------------------------------------------------- log marginal
@pytest.mark.parametrize("case", [0, 1])
def test_log_marginal_matches_mpmath(oracle, golden, case):
    c = golden["iterations"][case]
    S0, S5 = np.array(c["S0"]), np.array(c["exact"]["S5"])
    assert oracle.log_marginal_S(S0, c["alpha"], c["beta"], c["a"], c["b"]) == pytest.approx(
        c["log_marginal_S0"], rel=1e-13)
    assert oracle.log_marginal_S(S5, c["alpha"], c["beta"], c["a"], c["b"]) == pytest.approx(
        c["log_marginal_S5"], rel=1e-13)


def test_log_marginal_prefers_generating_params(oracle):
    # test_alloc_model_funcs.cpp:211-253 on a sample drawn from the model
    import bnmf_hip
    x, y, z = 8, 10, 2
    p = bnmf_hip.Params(40.0, 1.0, [1.0] * x, [1.0] * z)
    wins = 0
    for seed in range(10):
        S = bnmf_hip.sample_S(seed, (x, y, z), p)
        orig = oracle.log_marginal_S(S, p.alpha, p.beta, p.a, p.b)
        alt1 = oracle.log_marginal_S(S, p.alpha, p.beta, 200.0, 10.0)
        alt2 = oracle.log_marginal_S(S, [0.2] * x, [5.0] * z, 100.0, 1.0)
        wins += (orig > alt1) and (orig > alt2)
    assert wins == 10


# ---------------------------------------------------------------- sparseness
def test_sparseness_reference_cases(oracle):
    # test/util/test_util.cpp:31-55
    assert oracle.sparseness(np.zeros((5, 5, 5))) == np.finfo(float).max
    S = np.zeros((5, 5, 5))
    S[0, 0, 0] = 240
    assert oracle.sparseness(S) == pytest.approx(1.0)
    assert abs(oracle.sparseness(np.ones((5, 5, 5)))) <= np.finfo(float).eps


# ------------------------------------------------------------------ bld_fact
def test_bld_fact_is_the_posterior_mode(oracle):
    """bld_fact (bld_fact.hpp:36-80) returns the modes of the allocation
    model's conditional posteriors: W[:, k] ~ Dirichlet(alpha + S_i+k),
    H[:, j] ~ Dirichlet(beta + S_+jk), L_j ~ Gamma(a + S_+j+, rate b + 1).
    Each mode is found here by numerical maximisation of the log density
    (softmax parameterisation of the simplex, scipy.optimize), independently
    of the closed form, and compared with the oracle's factors; columns of W
    and H sum to 1."""
    from scipy.optimize import minimize, minimize_scalar
    from scipy.stats import gamma
    rng = np.random.default_rng(11)
    x, y, z = 5, 4, 3
    S = rng.uniform(0, 3, (x, y, z))
    alpha = rng.uniform(1.2, 2.5, x)
    beta = rng.uniform(1.2, 3.0, z)
    a, b = 4.0, 0.7
    W, H, L = oracle.bld_fact(S, alpha, beta, a, b)

    def dirichlet_mode(conc):
        # maximise sum (c_i - 1) log w_i over the simplex, w = softmax(u)
        def neg(u):
            lw = u - np.logaddexp.reduce(u)
            return -np.sum((conc - 1.0) * lw)
        res = minimize(neg, np.zeros(len(conc)), method="BFGS", options={"gtol": 1e-12})
        return np.exp(res.x - np.logaddexp.reduce(res.x))

    for k in range(z):
        assert np.allclose(W[:, k], dirichlet_mode(alpha + S[:, :, k].sum(1)), atol=1e-7), k
    for j in range(y):
        assert np.allclose(H[:, j], dirichlet_mode(beta + S[:, j, :].sum(0)), atol=1e-7), j
    for j in range(y):
        shape = a + S[:, j, :].sum()
        r = minimize_scalar(lambda v: -gamma.logpdf(v, shape, scale=1.0 / (b + 1.0)),
                            bounds=(1e-9, 10 * shape), method="bounded", options={"xatol": 1e-12})
        assert abs(L[j] - r.x) <= 1e-6 * r.x, j
    assert np.allclose(W.sum(0), 1.0, rtol=1e-14) and np.allclose(H.sum(0), 1.0, rtol=1e-14)


def test_bld_fact_closed_form(oracle):
    rng = np.random.default_rng(4)
    S = rng.uniform(0, 3, (4, 5, 3))
    alpha, beta, a, b = np.full(4, 1.5), np.array([1.0, 2.0, 3.0]), 5.0, 1.0
    W, H, L = oracle.bld_fact(S, alpha, beta, a, b)
    Wr = alpha[:, None] + S.sum(1) - 1
    Hr = beta[:, None] + S.sum(0).T - 1
    assert np.allclose(W, Wr / Wr.sum(0, keepdims=True), rtol=1e-14)
    assert np.allclose(H, Hr / Hr.sum(0, keepdims=True), rtol=1e-14)
    assert np.allclose(L, (a + S.sum(axis=(0, 2)) - 1) / (b + 1), rtol=1e-14)
