"""Solves at tile heights other than powers of two.

The tile table accepts any height from 4 to 1024 rows (the shipped table
holds 28 for the 1024x1024x32 FP64 class, and a tuner may write other
values). A table entry is a cap: an instance of fewer than 2^20 fibers gets
ceil(x / cap) row tiles of equal height, rounded up to a multiple of 4 (x =
130 under a cap of 40: four tiles of 36 rows, the last one 22). Each case sets
its shape's class to one cap, solves, and checks the result against the FP64
oracle: 1e-10 relative normwise for the plain FP64 form, 1e-4 for the
FP32-bulk form. The entries are put back afterwards.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("rows", [4, 12, 24, 28, 40, 100])
@pytest.mark.parametrize("dtype,x,y,z,appr", [(np.float64, 130, 70, 17, False),
                                              (np.float64, 101, 66, 32, False),
                                              (np.float32, 130, 70, 17, True),
                                              (np.float32, 205, 65, 7, True),
                                              (np.float32, 101, 66, 32, True)])
def test_odd_tile_heights_match_oracle(hip, oracle, rows, dtype, x, y, z, appr):
    iters = 40
    p = hip.Params(40.0, 1.0, [1.0] * x, [1.0] * z)
    X = hip.sample_X(7 * x + z, (x, y, z), p)
    S0 = hip.init_S(7 * x + z + 1, X, z).astype(dtype)
    old = hip.tune_get(dtype, z, x * y)
    hip.tune_set(dtype, z, x * y, rows)
    try:
        n_tiles = -(-x // rows)
        assert hip.tile_rows(dtype, x, y, z) == -(-(-(-x // n_tiles)) // 4) * 4
        S = hip.bld_mult(X.astype(dtype), z, p, iters, appr, S0=S0, dtype=dtype)
    finally:
        hip.tune_set(dtype, z, x * y, old)
    ref = oracle.bld_mult(X, z, p.alpha, p.beta, S0.astype(np.float64), max_iter=iters,
                          use_psi_appr=appr)
    tol = 1e-10 if dtype == np.float64 else 1e-4
    assert _rel(S, ref) <= tol, _rel(S, ref)


@pytest.mark.parametrize("rows", [4, 12, 28, 40, 100])
@pytest.mark.parametrize("dtype,x,y,z", [(np.float64, 130, 70, 17), (np.float32, 130, 70, 17),
                                         (np.float32, 101, 66, 32)])
def test_add_at_odd_tile_heights_matches_oracle(hip, oracle, rows, dtype, x, y, z):
    # bld_add runs the same step pass in its additive mode (tolerances as for
    # bld_add elsewhere: 1e-10 FP64, 1e-5 FP32)
    iters = 40
    p = hip.Params(40.0, 1.0, [1.0] * x, [1.0] * z)
    X = hip.sample_X(5 * x + z, (x, y, z), p)
    S0 = hip.init_S(5 * x + z + 1, X, z).astype(dtype)
    old = hip.tune_get(dtype, z, x * y)
    hip.tune_set(dtype, z, x * y, rows)
    try:
        S = hip.bld_add(X.astype(dtype), z, p, iters, S0=S0, dtype=dtype)
    finally:
        hip.tune_set(dtype, z, x * y, old)
    ref = oracle.bld_add(X, z, p.alpha, p.beta, S0.astype(np.float64), max_iter=iters)
    tol = 1e-10 if dtype == np.float64 else 1e-5
    assert _rel(S, ref) <= tol, _rel(S, ref)
