"""Autotuner on the device: it times candidates, keeps the fastest in the
tile table, and a solve planned with the tuned tiles still matches the FP64
oracle and is identical alone and batched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore_table(hip):
    # the test starts from an empty table (the library loads the shipped one)
    # and puts the loaded entries back afterwards
    snap = {(dt, z, f): hip.tune_get(dt, z, f) for dt in (np.float32, np.float64)
            for z in hip.TUNE_Z_CLASSES for f in (1 << 12, 1 << 15, 1 << 17, 1 << 19, 1 << 21)}
    for dt, z, f in snap:
        hip.tune_set(dt, z, f, 0)
    yield hip
    for (dt, z, f), rows in snap.items():
        hip.tune_set(dt, z, f, rows)


def test_autotune_small_classes(restore_table, oracle, tmp_path):
    hip = restore_table
    res = hip.autotune(np.float32, z_classes=(8, 24), fiber_classes=((1 << 10, 1 << 14),),
                       candidates=(16, 32, 64), reps=2, target_elements=1 << 20,
                       path=str(tmp_path / "t.txt"))
    assert set(res) == {(1, 0), (4, 0)}   # z classes ending at 8 and 24, x*y < 2^14
    for (zc, fc), r in res.items():
        assert r["rows"] in (16, 32, 64)
        assert r["pass_ms"][r["rows"]] == min(r["pass_ms"].values())
        side, _, z = r["shape"]
        assert hip.tune_get(np.float32, z, side * side) == r["rows"]
    assert (tmp_path / "t.txt").read_text().count("f32") == 2
    # a tuned shape: FP32 within 1e-4 of the FP64 oracle, same bits alone and batched
    side, _, z = res[(4, 0)]["shape"]
    rng = np.random.default_rng(3)
    p = hip.Params(40.0, 1.0, [1.0] * side, [1.0] * z)
    X = hip.sample_X(11, (side, side, z), p)
    S0 = hip.init_S(12, X, z)
    ref = oracle.bld_mult(X, z, p.alpha, p.beta, S0, max_iter=60, use_psi_appr=True)
    inst = hip.Instance(X=X.astype(np.float32), z=z, params=p, S0=S0.astype(np.float32))
    other = hip.Instance(X=rng.poisson(3.0, (40, 70)).astype(np.float32), z=5,
                         params=hip.Params(40.0, 1.0, [1.0] * 40, [1.0] * 5), seed=4)
    with hip.Batch([inst], dtype=np.float32) as a, hip.Batch([other, inst], dtype=np.float32) as b:
        a.solve(60)
        b.solve(60)
        Sa, Sb = a.S(0), b.S(1)
    assert np.array_equal(Sa, Sb)
    err = np.linalg.norm(Sa - ref) / np.linalg.norm(ref)
    assert err <= 1e-4, err
