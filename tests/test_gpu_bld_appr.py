"""GPU parity of the approximate solver (bld::bld_appr,
src/bld/bld_appr.hpp:287-372) against the oracle's restatement (pinned in
tests/test_oracle_bld_appr.py).

Tolerances (per test):
  * FP64 vs the oracle: 1e-12 relative normwise on nu and mu after one
    iteration, 1e-10 after whole solves (measured on MI355X: up to 4e-13
    after 100 iterations), and the same on S;
  * FP32 vs the FP64 oracle: 5e-5 relative normwise on S after 100
    iterations (measured: 3.5e-7 to 4.4e-6);
  * batched vs alone: bitwise.
"""
import numpy as np
import pytest

from test_oracle_bld_appr import APPR_FIXTURE_X

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)


class _Inst:
    def __init__(self, X, z, params, nu0, mu0):
        self.X, self.z, self.params, self.nu0, self.mu0, self.seed = X, z, params, nu0, mu0, 0


def _problem(hip, x, y, z, seed):
    p = hip.Params(40.0, 1.0, [1.0] * x, [1.0] * z)
    X = hip.sample_X(seed, (x, y, z), p)
    rng = np.random.default_rng(seed)
    nu0 = rng.dirichlet(np.ones(z), size=x)
    mu0 = rng.dirichlet(np.ones(z), size=y).T.copy()
    return p, X, nu0, mu0


SHAPES = [(4, 5, 3), (37, 70, 5), (70, 130, 17), (20, 65, 64), (129, 64, 8), (50, 40, 31)]


@pytest.mark.parametrize("shape", SHAPES)
def test_appr_fp64_one_iteration_matches_oracle(hip, oracle, shape):
    x, y, z = shape
    p, X, nu0, mu0 = _problem(hip, x, y, z, 7)
    S, nu, mu = hip.bld_appr(X, z, p, 1, nu0=nu0, mu0=mu0)
    Sr, nur, mur = oracle.bld_appr(X, z, p.alpha, p.beta, nu0, mu0, max_iter=1)
    assert _rel(nu, nur) < 1e-12 and _rel(mu, mur) < 1e-12, (_rel(nu, nur), _rel(mu, mur))
    assert _rel(S, Sr) < 1e-12


@pytest.mark.parametrize("shape", SHAPES)
def test_appr_fp64_solve_matches_oracle(hip, oracle, shape):
    x, y, z = shape
    p, X, nu0, mu0 = _problem(hip, x, y, z, 3)
    iters = 60 if x * y * z < 200000 else 25
    S, nu, mu = hip.bld_appr(X, z, p, iters, nu0=nu0, mu0=mu0)
    Sr, nur, mur = oracle.bld_appr(X, z, p.alpha, p.beta, nu0, mu0, max_iter=iters)
    assert _rel(nu, nur) < 1e-10 and _rel(mu, mur) < 1e-10, (_rel(nu, nur), _rel(mu, mur))
    assert _rel(S, Sr) < 1e-10
    assert np.abs(S.sum(2) - X).max() <= 1e-12 * max(X.max(), 1)


def test_appr_fixture_known_answer(hip, oracle):
    """test/bld/test_bld_algs.cpp:433-471: log-marginal >= -280, sparseness
    >= 0.50 and S_ij+ = X after the default 1000 iterations."""
    X = APPR_FIXTURE_X
    p = hip.Params(40, 1, [1.0] * 8, [1.0] * 3)
    rng = np.random.default_rng(0)
    nu0 = rng.dirichlet(np.ones(3), size=8)
    mu0 = rng.dirichlet(np.ones(3), size=10).T.copy()
    S, nu, mu = hip.bld_appr(X, 3, p, nu0=nu0, mu0=mu0)
    Sr, _, _ = oracle.bld_appr(X, 3, p.alpha, p.beta, nu0, mu0)
    assert _rel(S, Sr) < 1e-10
    assert hip.log_marginal_S(S, p) >= -280
    assert oracle.sparseness(S) >= 0.50
    assert np.linalg.norm(S.sum(2) - X) / np.linalg.norm(X) < 1e-12
    # the seeded draw path (no explicit factors) also runs and keeps S_ij+ = X
    S2, nu2, mu2 = hip.bld_appr(X, 3, p, 50, seed=5)
    assert np.allclose(nu2.sum(1), 1, atol=1e-12) and np.allclose(mu2.sum(0), 1, atol=1e-12)
    assert np.abs(S2.sum(2) - X).max() < 1e-12 * X.max()


def test_appr_zero_matrix(hip):
    p = hip.Params(1, 1, [1.0] * 8, [1.0] * 3)
    for dt in (np.float64, np.float32):
        S, nu, mu = hip.bld_appr(np.zeros((8, 10)), 3, p, 30, seed=1, dtype=dt)
        assert np.all(S == 0) and np.all(np.isfinite(nu)) and np.all(np.isfinite(mu))


@pytest.mark.parametrize("shape", [(37, 70, 5), (70, 130, 17), (20, 65, 64), (50, 40, 31)])
def test_appr_fp32_close_to_fp64_oracle(hip, oracle, shape):
    x, y, z = shape
    p, X, nu0, mu0 = _problem(hip, x, y, z, 5)
    nu0f, mu0f = nu0.astype(np.float32), mu0.astype(np.float32)
    S, nu, mu = hip.bld_appr(X.astype(np.float32), z, p, 100, nu0=nu0f, mu0=mu0f, dtype=np.float32)
    Sr, _, _ = oracle.bld_appr(X, z, p.alpha, p.beta, nu0f.astype(np.float64),
                               mu0f.astype(np.float64), max_iter=100)
    assert _rel(S, Sr) < 5e-5, _rel(S, Sr)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_appr_batch_matches_alone_bitwise(hip, dtype):
    insts, alone = [], []
    for n, (x, y, z) in enumerate(SHAPES):
        p, X, nu0, mu0 = _problem(hip, x, y, z, 30 + n)
        X, nu0, mu0 = X.astype(dtype), nu0.astype(dtype), mu0.astype(dtype)
        insts.append(_Inst(X, z, p, nu0, mu0))
        alone.append(hip.bld_appr(X, z, p, 12, nu0=nu0, mu0=mu0, dtype=dtype))
    with hip.ApprBatch(insts, dtype=dtype) as b:
        b.solve(12)
        for n in range(len(SHAPES)):
            got = b.get(n)
            for g, a in zip(got, alone[n]):
                assert np.array_equal(g, a), SHAPES[n]
