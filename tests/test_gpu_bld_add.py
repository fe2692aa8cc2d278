"""GPU parity of the additive solver (bld::bld_add, src/bld/bld_add.hpp:69-184)
against the oracle's restatement (pinned in tests/test_oracle_bld_add.py).

Tolerances (per test):
  * FP64 form vs the oracle: 1e-12 relative normwise after one and two steps
    (the step size depends on the iteration number, so these pin its
    indexing), 1e-10 after whole solves;
  * FP32 form vs the FP64 oracle: 1e-5 relative normwise after 200 steps
    (measured on MI355X: 5e-7 to 1e-6 on these shapes);
  * batched vs alone: bitwise; a second solve of a batch repeats the first
    bitwise (the step counter restarts).
"""
import numpy as np
import pytest

from test_oracle_bld_add import ADD_FIXTURE_X

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)


def _problem(hip, x, y, z, seed, a=40.0):
    p = hip.Params(a, 1.0, [1.0] * x, [1.0] * z)
    X = hip.sample_X(seed, (x, y, z), p)
    S0 = hip.init_S(seed + 1, X, z)
    return p, X, S0


@pytest.mark.parametrize("shape", [(3, 4, 3), (37, 70, 5), (70, 130, 17), (20, 65, 64)])
def test_add_fp64_first_steps_match_oracle(hip, oracle, shape):
    x, y, z = shape
    p, X, S0 = _problem(hip, x, y, z, 7)
    for iters in (1, 2):
        got = hip.bld_add(X, z, p, iters, S0=S0)
        ref = oracle.bld_add(X, z, p.alpha, p.beta, S0, max_iter=iters)
        assert _rel(got, ref) < 1e-12, (shape, iters, _rel(got, ref))


@pytest.mark.parametrize("shape,iters", [((37, 70, 5), 100), ((70, 130, 17), 60),
                                         ((129, 64, 8), 60), ((20, 65, 64), 40),
                                         ((50, 40, 31), 60)])
def test_add_fp64_solve_matches_oracle(hip, oracle, shape, iters):
    x, y, z = shape
    p, X, S0 = _problem(hip, x, y, z, 3)
    got = hip.bld_add(X, z, p, iters, S0=S0)
    ref = oracle.bld_add(X, z, p.alpha, p.beta, S0, max_iter=iters)
    assert _rel(got, ref) < 1e-10, _rel(got, ref)
    assert np.abs(got.sum(2) - X).max() <= 1e-12 * max(X.max(), 1)


def test_add_fixture_known_answer(hip, oracle):
    """test/bld/test_bld_algs.cpp:248-289 from the init_S draw of seed 1
    (a start from which the reference's thresholds hold)."""
    X = ADD_FIXTURE_X
    p = hip.Params(40, 1, [1.0] * 8, [1.0] * 3)
    S0 = hip.init_S(1, X, 3)
    S = hip.bld_add(X, 3, p, S0=S0)
    ref = oracle.bld_add(X, 3, p.alpha, p.beta, S0)
    assert _rel(S, ref) < 1e-10
    assert hip.log_marginal_S(S, p) >= -273
    assert oracle.sparseness(S) >= 0.58
    assert np.linalg.norm(S.sum(2) - X) / np.linalg.norm(X) < 1e-12


def test_add_zero_and_degenerate_inputs(hip):
    p = hip.Params(1, 1, [1.0] * 8, [1.0] * 3)
    S = hip.bld_add(np.zeros((8, 10)), 3, p, 50)
    assert np.all(S == 0)
    Sf = hip.bld_add(np.zeros((8, 10)), 3, p, 50, dtype=np.float32)
    assert np.all(Sf == 0)
    rng = np.random.default_rng(2)
    X = rng.integers(0, 6, (9, 70)).astype(float)
    X[2, :] = 0
    X[:, 65] = 0
    p9 = hip.Params(1, 1, [1.0] * 9, [1.0] * 4)
    for dt in (np.float64, np.float32):
        S = hip.bld_add(X, 4, p9, 30, dtype=dt)
        assert np.all(np.isfinite(S)) and np.all(S[X == 0] == 0)
    p1 = hip.Params(1, 1, [1.0] * 9, [1.0])
    S1 = hip.bld_add(X, 1, p1, 20)
    assert np.allclose(S1[:, :, 0], X, rtol=1e-14, atol=0)


@pytest.mark.parametrize("beta", [1.0, 0.5])
@pytest.mark.parametrize("shape", [(37, 70, 5), (70, 130, 17), (20, 65, 64), (50, 40, 31)])
def test_add_fp32_close_to_fp64_oracle(hip, oracle, shape, beta):
    # beta >= 1 runs the exact-z FP32 kernel, beta < 1 the generic one
    x, y, z = shape
    p, X, S0 = _problem(hip, x, y, z, 5)
    p = hip.Params(p.a, p.b, p.alpha, [beta] * z)
    got = hip.bld_add(X.astype(np.float32), z, p, 200, S0=S0.astype(np.float32), dtype=np.float32)
    ref = oracle.bld_add(X, z, p.alpha, p.beta, S0.astype(np.float32).astype(np.float64),
                         max_iter=200)
    assert _rel(got, ref) < 1e-5, _rel(got, ref)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_add_batch_matches_alone_bitwise(hip, dtype):
    shapes = [(37, 70, 5), (70, 130, 17), (20, 65, 64), (129, 64, 8), (50, 40, 31), (8, 10, 3)]
    insts, alone = [], []
    for n, (x, y, z) in enumerate(shapes):
        p, X, S0 = _problem(hip, x, y, z, 20 + n)
        X, S0 = X.astype(dtype), S0.astype(dtype)
        insts.append(hip.Instance(X, z, p, S0=S0))
        alone.append(hip.bld_add(X, z, p, 25, S0=S0, dtype=dtype))
    with hip.Batch(insts, dtype=dtype, algo=hip.ALGO_ADD) as b:
        b.solve(25)
        first = [b.S(n) for n in range(len(shapes))]
        b.solve(25)
        for n in range(len(shapes)):
            assert np.array_equal(first[n], alone[n]), shapes[n]
            assert np.array_equal(b.S(n), first[n]), shapes[n]
