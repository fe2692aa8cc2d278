"""Pins the oracle's restatement of bld::bld_add (oracle/bld_oracle.cpp,
iterate_add; reference src/bld/bld_add.hpp:69-184) before it is trusted.

Checks independent of the restated formula:
  * grad_S is d log p(S) / dS_ijk (central differences of log_marginal_S,
    itself pinned against mpmath) up to a term constant along k;
  * the step is Z += eta * X * dlog p / dZ in the softmax parameterisation
    S = X softmax_k(Z), with dlog p / dZ formed from the numerical gradient
    and eta(step) = 0.1 / (step + 1)^0.55 (bld_add.hpp:101-103);
  * S_ij+ = X after every step; X = 0 gives S = 0; z = 1 gives S = X;
  * the reference's known-answer case (test/bld/test_bld_algs.cpp:248-289).
"""
import numpy as np
import pytest


def _num_grad(oracle, S, alpha, beta, a, b, h=1e-5):
    d = np.empty_like(S)
    for idx in np.ndindex(*S.shape):
        Sp, Sm = S.copy(), S.copy()
        Sp[idx] += h
        Sm[idx] -= h
        d[idx] = (oracle.log_marginal_S(Sp, alpha, beta, a, b) -
                  oracle.log_marginal_S(Sm, alpha, beta, a, b)) / (2 * h)
    return d


def _case(seed=5):
    rng = np.random.default_rng(seed)
    x, y, z = 3, 4, 3
    alpha = np.array([0.5, 1.0, 2.0])
    beta = np.array([1.0, 3.0, 0.7])
    S = rng.uniform(0.5, 4.0, (x, y, z))
    return S.sum(2), alpha, beta, S


def test_add_gradient_is_log_marginal_gradient(oracle):
    X, alpha, beta, S = _case()
    a, b = 7.0, 2.0
    _, G = oracle.bld_add_step(X, alpha, beta, S, eph=0)
    resid = G - _num_grad(oracle, S, alpha, beta, a, b)
    # psi_appr is within 1e-10 of psi; central differences within ~1e-8
    assert np.abs(resid - resid.mean(axis=2, keepdims=True)).max() < 1e-7


@pytest.mark.parametrize("eph", [0, 7, 999])
def test_add_step_is_softmax_ascent(oracle, eph):
    X, alpha, beta, S = _case(11)
    S1, _ = oracle.bld_add_step(X, alpha, beta, S, eph=eph)
    g = _num_grad(oracle, S, alpha, beta, 7.0, 2.0)
    # d log p / dZ_k = (S_k / X) (X g_k - sum_m S_m g_m) for S = X softmax(Z);
    # a term of g constant along k drops out because sum_k S_k = X
    D = S * (X[:, :, None] * g - (S * g).sum(2, keepdims=True))
    eta = 0.1 / (eph + 1) ** 0.55
    lhs = np.log(S1 / S1[:, :, :1]) - np.log(S / S[:, :, :1])
    rhs = eta * (D - D[:, :, :1])
    assert np.abs(lhs - rhs).max() < 1e-6 * max(1.0, np.abs(rhs).max())
    assert np.abs(S1.sum(2) - X).max() < 1e-12 * X.max()


def test_add_step_raises_log_marginal_for_small_steps(oracle):
    X, alpha, beta, S = _case(3)
    S1, _ = oracle.bld_add_step(X, alpha, beta, S, eph=100000)
    assert (oracle.log_marginal_S(S1, alpha, beta, 7.0, 2.0) >
            oracle.log_marginal_S(S, alpha, beta, 7.0, 2.0))


def test_add_zero_matrix_gives_zero(oracle):
    X = np.zeros((8, 10))
    S0 = np.zeros((8, 10, 3))
    S = oracle.bld_add(X, 3, np.ones(8), np.ones(3), S0, max_iter=20)
    assert np.all(S == 0)


def test_add_zero_fibers_stay_zero(oracle):
    rng = np.random.default_rng(2)
    X = rng.integers(0, 6, (7, 9)).astype(float)
    X[2, :] = 0
    X[:, 4] = 0
    S0 = X[:, :, None] * rng.dirichlet(np.ones(4), size=(7, 9))
    S = oracle.bld_add(X, 4, np.ones(7), np.ones(4), S0, max_iter=30)
    assert np.all(S[X == 0] == 0) and np.all(np.isfinite(S))
    assert np.abs(S.sum(2) - X).max() < 1e-12 * X.max()


def test_add_z1_gives_X(oracle):
    rng = np.random.default_rng(4)
    X = rng.integers(0, 9, (6, 5)).astype(float)
    S = oracle.bld_add(X, 1, np.ones(6), np.ones(1), X[:, :, None].copy(), max_iter=40)
    assert np.allclose(S[:, :, 0], X, rtol=1e-14, atol=0)


def test_add_rejects_negative_X(oracle):
    X = np.full((4, 3), 5.0)
    X[0, 0] = -5
    with pytest.raises(ValueError):
        oracle.bld_add(X, 2, np.ones(4), np.ones(2), np.ones((4, 3, 2)), max_iter=5)


# test/bld/test_bld_algs.cpp:248-289 (Experiments.ipynb data, z = 3,
# Params(40, 1, 1_x, 1_z), default 1000 iterations)
ADD_FIXTURE_X = np.array([
    6., 3., 7., 15., 8., 6., 1., 7., 6., 7., 3., 3., 11., 1., 3., 3., 0., 1., 3., 1., 1., 2., 2.,
    8., 3., 2., 0., 2., 6., 1., 8., 9., 18., 5., 8., 6., 19., 6., 11., 8., 6., 3., 5., 1., 2., 0.,
    2., 4., 0., 2., 2., 1., 2., 2., 1., 0., 1., 2., 7., 1., 6., 10., 15., 6., 3., 2., 9., 4., 12.,
    4., 3., 2., 0., 4., 2., 2., 1., 3., 2., 2.]).reshape(8, 10)


def test_add_fixture_known_answer(oracle):
    """The reference's thresholds, from Dirichlet(1) starts; like the
    multiplicative case they hold from most starting points but not all, so
    the check is on a majority of eight seeded starts."""
    X = ADD_FIXTURE_X
    alpha, beta = np.ones(8), np.ones(3)
    ok = 0
    for seed in range(8):
        S0 = X[:, :, None] * np.random.default_rng(seed).dirichlet(np.ones(3), size=(8, 10))
        S = oracle.bld_add(X, 3, alpha, beta, S0)
        lm = oracle.log_marginal_S(S, alpha, beta, 40, 1)
        sp = oracle.sparseness(S)
        gap = np.linalg.norm(S.sum(2) - X) / np.linalg.norm(X)
        assert gap < 1e-12
        assert lm > -285 and sp > 0.55
        ok += lm >= -273 and sp >= 0.58
    assert ok >= 4
