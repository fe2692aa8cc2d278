"""Degenerate shapes through every solver and precision against the oracle:
a single row, a single column, a single fiber, z = 1 and z = 64 on a 2 x 2
matrix (ragged tiles everywhere). FP64 to 1e-10 relative, FP32 to 1e-4
(few iterations: indexing mistakes show at once)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (1, 1, 5), (1, 7, 3), (7, 1, 3), (2, 2, 64), (1, 70, 33), (65, 1, 17),
          (3, 65, 4)]


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)


def _case(hip, x, y, z, seed):
    p = hip.Params(40.0, 1.0, [1.0] * x, [1.0] * z)
    rng = np.random.default_rng(seed)
    X = rng.integers(0, 9, (x, y)).astype(np.float64)
    X.flat[0] = max(X.flat[0], 1.0)
    S0 = hip.init_S(seed, X, z)
    return p, X, S0


@pytest.mark.parametrize("shape", SHAPES)
def test_edges_bld_mult(hip, oracle, shape):
    x, y, z = shape
    p, X, S0 = _case(hip, x, y, z, 3)
    for appr in (False, True):
        S = hip.bld_mult(X, z, p, 6, appr, S0=S0)
        ref = oracle.bld_mult(X, z, p.alpha, p.beta, S0, max_iter=6, use_psi_appr=appr)
        assert np.all(np.isfinite(S)) and _rel(S, ref) < 1e-10, (shape, appr, _rel(S, ref))
    S32 = hip.bld_mult(X.astype(np.float32), z, p, 6, True, S0=S0.astype(np.float32),
                       dtype=np.float32)
    ref32 = oracle.bld_mult(X, z, p.alpha, p.beta, S0.astype(np.float32).astype(np.float64),
                            max_iter=6, use_psi_appr=True)
    assert np.all(np.isfinite(S32)) and _rel(S32, ref32) < 1e-4, (shape, _rel(S32, ref32))


@pytest.mark.parametrize("shape", SHAPES)
def test_edges_bld_add(hip, oracle, shape):
    x, y, z = shape
    p, X, S0 = _case(hip, x, y, z, 4)
    S = hip.bld_add(X, z, p, 6, S0=S0)
    ref = oracle.bld_add(X, z, p.alpha, p.beta, S0, max_iter=6)
    assert np.all(np.isfinite(S)) and _rel(S, ref) < 1e-10, (shape, _rel(S, ref))
    S32 = hip.bld_add(X.astype(np.float32), z, p, 6, S0=S0.astype(np.float32), dtype=np.float32)
    assert np.all(np.isfinite(S32)) and _rel(S32, ref) < 1e-4, (shape, _rel(S32, ref))


# bld_appr with a single row degenerates in the reference's own arithmetic
# (grad_plus = grad_minus = 0 gives nu = 0 after the first half-step, then
# X / (0 + eps) ~ 1e50 overflows): not a meaningful comparison
@pytest.mark.parametrize("shape", [s for s in SHAPES if s[0] > 1])
def test_edges_bld_appr(hip, oracle, shape):
    x, y, z = shape
    p, X, _ = _case(hip, x, y, z, 5)
    rng = np.random.default_rng(6)
    nu0 = rng.dirichlet(np.ones(z), size=x)
    mu0 = rng.dirichlet(np.ones(z), size=y).T.copy()
    S, nu, mu = hip.bld_appr(X, z, p, 6, nu0=nu0, mu0=mu0)
    Sr, nur, mur = oracle.bld_appr(X, z, p.alpha, p.beta, nu0, mu0, max_iter=6)
    assert np.all(np.isfinite(S)) and _rel(S, Sr) < 1e-10, (shape, _rel(S, Sr))
    S32, _, _ = hip.bld_appr(X.astype(np.float32), z, p, 6, nu0=nu0.astype(np.float32),
                             mu0=mu0.astype(np.float32), dtype=np.float32)
    assert np.all(np.isfinite(S32)) and _rel(S32, Sr) < 1e-4, (shape, _rel(S32, Sr))


@pytest.mark.parametrize("shape", [(0, 5, 3), (4, 0, 3), (0, 0, 2)])
def test_empty_inputs(hip, shape):
    x, y, z = shape
    p = hip.Params(1.0, 1.0, [1.0] * x, [1.0] * z)
    X = np.zeros((x, y))
    for dt in (np.float64, np.float32):
        assert hip.bld_mult(X, z, p, 5, dtype=dt).shape == (x, y, z)
        assert hip.bld_add(X, z, p, 5, dtype=dt).shape == (x, y, z)
        S, nu, mu = hip.bld_appr(X, z, p, 5, dtype=dt)
        assert S.shape == (x, y, z) and nu.shape == (x, z) and mu.shape == (z, y)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_empty_instance_in_batches(hip, dtype):
    # an empty instance among others: the others solve as they do alone
    pa, Xa, S0a = _case(hip, 20, 30, 5, 8)
    pe = hip.Params(1.0, 1.0, [], [1.0] * 4)
    insts = [hip.Instance(X=Xa.astype(dtype), z=5, params=pa, S0=S0a.astype(dtype)),
             hip.Instance(X=np.zeros((0, 7), dtype=dtype), z=4, params=pe)]
    alone = hip.bld_mult(Xa.astype(dtype), 5, pa, 10, True, S0=S0a.astype(dtype), dtype=dtype)
    with hip.Batch(insts, dtype=dtype) as b:
        b.solve(10)
        assert np.array_equal(b.S(0), alone)
        assert b.S(1).shape == (0, 7, 4)
        lm = b.log_marginal()
        assert np.isfinite(lm[0])

    class _I:
        pass
    ia, ie = _I(), _I()
    ia.X, ia.z, ia.params, ia.nu0, ia.mu0, ia.seed = Xa.astype(dtype), 5, pa, None, None, 3
    ie.X, ie.z, ie.params, ie.nu0, ie.mu0, ie.seed = np.zeros((0, 7), dtype=dtype), 4, pe, None, None, 4
    one = hip.bld_appr(Xa.astype(dtype), 5, pa, 10, seed=3, dtype=dtype)
    with hip.ApprBatch([ia, ie], dtype=dtype) as b:
        b.solve(10)
        got = b.get(0)
        for g, o in zip(got, one):
            assert np.array_equal(g, o)
        S, nu, mu = b.get(1)
        assert S.shape == (0, 7, 4) and mu.shape == (4, 7)
        # no iteration runs on it: mu stays the seeded Dirichlet draw, alone
        # and batched
        _, _, mu_alone = hip.bld_appr(np.zeros((0, 7), dtype=dtype), 4, pe, 10, seed=4, dtype=dtype)
        assert np.array_equal(mu, mu_alone)
        assert np.allclose(mu.sum(0), 1.0, atol=1e-6)
