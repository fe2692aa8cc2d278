"""Pins the oracle's restatement of bld::bld_appr (oracle/bld_oracle.cpp,
appr_*; reference src/bld/bld_appr.hpp) before it is trusted.

Checks independent of the restated formulas:
  * with S(nu, mu)_ijk = X_ij nu_ik mu_kj / (nu mu)_ij, the numerator minus
    the denominator of each multiplicative update is the gradient of
    log p(S(nu, mu)) (central differences of log_marginal_S, itself pinned
    against mpmath) with respect to nu (update_nu) or mu (update_mu): the
    update is the usual split of that gradient into its positive and
    negative parts;
  * a half-step multiplies nu_ik by nom/den and renormalises the rows of nu
    (columns of mu) to sum to 1;
  * S_ij+ = X; X = 0 gives S = 0;
  * the reference's known-answer case (test/bld/test_bld_algs.cpp:433-471).
"""
import numpy as np
import pytest


def _S(X, nu, mu):
    h = nu @ mu
    return (X / h)[:, :, None] * nu[:, None, :] * mu.T[None, :, :]


def _case(seed):
    rng = np.random.default_rng(seed)
    x, y, z = 4, 5, 3
    alpha = np.array([0.5, 1.0, 2.0, 1.5])
    beta = np.array([1.0, 3.0, 0.7])
    X = rng.integers(1, 9, (x, y)).astype(float)
    nu = rng.dirichlet(np.ones(z), size=x)
    mu = rng.dirichlet(np.ones(z), size=y).T.copy()
    return X, alpha, beta, nu, mu


def _num_grad(oracle, X, alpha, beta, nu, mu, wrt, a=40.0, b=1.0, h=1e-6):
    P = nu if wrt == 0 else mu
    g = np.empty_like(P)
    for idx in np.ndindex(*P.shape):
        vals = []
        for sgn in (1, -1):
            Q = P.copy()
            Q[idx] += sgn * h
            n2, m2 = (Q, mu) if wrt == 0 else (nu, Q)
            vals.append(oracle.log_marginal_S(_S(X, n2, m2), alpha, beta, a, b))
        g[idx] = (vals[0] - vals[1]) / (2 * h)
    return g


@pytest.mark.parametrize("half", [0, 1])
@pytest.mark.parametrize("seed", [1, 2])
def test_appr_update_splits_the_gradient(oracle, half, seed):
    X, alpha, beta, nu, mu = _case(seed)
    out, nom, den = oracle.bld_appr_half(X, alpha, beta, nu, mu, half)
    g = _num_grad(oracle, X, alpha, beta, nu, mu, half)
    # psi_appr is within 1e-10 of psi, central differences within ~1e-7
    assert np.abs((nom - den) - g).max() < 1e-6 * max(1.0, np.abs(g).max())
    P = nu if half == 0 else mu
    step = P * nom / den
    if half == 0:
        expect = step / step.sum(1, keepdims=True)
    else:
        expect = step / step.sum(0, keepdims=True)
    assert np.abs(out - expect).max() < 1e-14


def test_appr_stationary_point_keeps_factors(oracle):
    """At a point where the gradient along the simplex vanishes the step is
    (close to) the identity: take many steps, then one more changes little."""
    X, alpha, beta, nu, mu = _case(3)
    S, nu1, mu1 = oracle.bld_appr(X, 3, alpha, beta, nu, mu, max_iter=400)
    S2, nu2, mu2 = oracle.bld_appr(X, 3, alpha, beta, nu1, mu1, max_iter=1)
    assert np.abs(nu2 - nu1).max() < 1e-3 and np.abs(mu2 - mu1).max() < 1e-3
    assert np.allclose(nu1.sum(1), 1, atol=1e-14) and np.allclose(mu1.sum(0), 1, atol=1e-14)


def test_appr_sums_and_zero(oracle):
    X, alpha, beta, nu, mu = _case(4)
    S, nu1, mu1 = oracle.bld_appr(X, 3, alpha, beta, nu, mu, max_iter=20)
    assert np.abs(S.sum(2) - X).max() < 1e-13 * X.max()
    assert np.allclose(S, _S(X, nu1, mu1), rtol=1e-14, atol=0)
    Z = np.zeros_like(X)
    S0, _, _ = oracle.bld_appr(Z, 3, alpha, beta, nu, mu, max_iter=20)
    assert np.all(S0 == 0)


def test_appr_rejects_negative_X(oracle):
    X, alpha, beta, nu, mu = _case(5)
    X[0, 0] = -5
    with pytest.raises(ValueError):
        oracle.bld_appr(X, 3, alpha, beta, nu, mu, max_iter=5)


# test/bld/test_bld_algs.cpp:433-471 (Experiments.ipynb data, z = 3,
# Params(40, 1, 1_x, 1_z), default 1000 iterations)
APPR_FIXTURE_X = np.array([
    2., 0., 2., 2., 0., 2., 1., 0., 1., 3., 5., 4., 0., 2., 2., 6., 2., 1., 4., 5., 1., 13., 6.,
    9., 5., 8., 6., 6., 11., 12., 6., 2., 1., 4., 9., 7., 6., 10., 4., 8., 5., 5., 4., 3., 1.,
    0., 4., 8., 3., 1., 0., 9., 2., 1., 2., 7., 5., 3., 8., 3., 9., 8., 4., 5., 11., 9., 4., 6.,
    5., 12., 2., 6., 3., 3., 3., 6., 5., 3., 6., 4.]).reshape(8, 10)


def test_appr_fixture_known_answer(oracle):
    X = APPR_FIXTURE_X
    alpha, beta = np.ones(8), np.ones(3)
    ok = 0
    for seed in range(6):
        rng = np.random.default_rng(seed)
        nu0 = rng.dirichlet(np.ones(3), size=8)
        mu0 = rng.dirichlet(np.ones(3), size=10).T.copy()
        S, nu, mu = oracle.bld_appr(X, 3, alpha, beta, nu0, mu0)
        assert np.linalg.norm(S.sum(2) - X) / np.linalg.norm(X) < 1e-12
        ok += (oracle.log_marginal_S(S, alpha, beta, 40, 1) >= -280 and
               oracle.sparseness(S) >= 0.50)
    assert ok >= 5
