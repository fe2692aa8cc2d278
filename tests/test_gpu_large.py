"""One instance beyond 2^31 elements (8200 x 4100 x 64 = 2.15e9; S is 8.6 GB
in FP32): the pass, prep and partial-sum indexing must be 64-bit throughout.
Checked through size-independent properties: the additive step keeps
S_ij+ = X after every iteration (normalised per fiber), S stays finite and
nonnegative, and a repeated solve is bitwise identical (the multiplicative
step, same kernel family). The initial tensor is drawn on the device.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_fp32_instance_beyond_2_31_elements(hip):
    x, y, z = 8200, 4100, 64
    assert x * y * z > 2**31
    p = hip.Params(40.0, 1.0, [1.0] * x, [1.0] * z)
    rng = np.random.default_rng(7)
    X = rng.integers(0, 40, (x, y)).astype(np.float32)
    inst = hip.Instance(X=X, z=z, params=p, seed=11)
    print(f"[large] bld_add {x}x{y}x{z} FP32", flush=True)
    with hip.Batch([inst], dtype=np.float32, algo=hip.ALGO_ADD) as b:
        b.solve(2)
        S = b.S(0)
    rowsum = S.sum(axis=2, dtype=np.float64)
    assert np.all(np.isfinite(rowsum))
    assert np.abs(rowsum - X).max() <= 2e-5 * X.max()
    assert S.min() >= 0.0
    # rows at the far end of the tensor (offsets beyond 2^31 elements)
    assert np.abs(rowsum[-3:] - X[-3:]).max() <= 2e-5 * X.max()
    del S, rowsum
    print(f"[large] bld_mult {x}x{y}x{z} FP32, twice", flush=True)
    with hip.Batch([inst], dtype=np.float32) as b:
        b.solve(2)
        S1 = b.S(0)
        b.solve(2)
        S2 = b.S(0)
    assert np.array_equal(S1, S2)
    assert np.isfinite(S1[-1]).all() and S1.min() >= 0.0
