"""Parity at the benchmark configurations' full sizes.

  * 1024 x 1024 x 32, 500 iterations, exact digamma (the plain FP64 form):
    S and log_marginal_S within 1e-10 relative of the FP64 CPU oracle on the
    same X, seed and initial tensor.
  * 2048 x 2048 x 64, 300 iterations, FP32 bulk with psi_appr: S within 1e-4
    normwise of the FP64 form on the same start, log-marginal gap <= 1e-5.
  * The batched headline workload (256 instances): eight of its instances
    against the FP64 oracle directly (FP32 within 1e-4, the FP64 form within
    1e-10), chosen as the four with the largest FP32-vs-FP64 gap and four of
    different depth z.
  * Serving: 48 requests drawn from the serving workload (x, y log-uniform in
    [64, 2048], z log-uniform in [8, 64]) submitted together; each result
    equals the same request solved alone bit for bit and lies within 1e-4 of
    its FP64 solve.
Inputs come from bench.py's generators, so the tests check exactly what the
benchmark times.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def test_fp64_1024x1024x32_500_iterations_matches_oracle(hip, oracle):
    import bench
    inst = bench.single(hip, 1024, 1024, 32, 40.0, 0xB1D1)
    p = inst.params
    print("[c1] GPU FP64 exact 1024x1024x32, 500 iterations", flush=True)
    S_gpu = hip.bld_mult(inst.X, 32, p, 500, False, S0=inst.S0)
    print("[c1] oracle", flush=True)
    S_cpu = oracle.bld_mult(inst.X, 32, p.alpha, p.beta, inst.S0, max_iter=500)
    err = _rel(S_gpu, S_cpu)
    print(f"[c1] rel err {err:.3e}", flush=True)
    assert err <= 1e-10
    lm_gpu = hip.log_marginal_S(S_gpu, p)
    lm_cpu = oracle.log_marginal_S(S_cpu, p.alpha, p.beta, p.a, p.b)
    assert abs(lm_gpu - lm_cpu) <= 1e-10 * abs(lm_cpu)


def test_fp32_2048x2048x64_300_iterations_within_1e4_of_fp64(hip):
    import bench
    inst = bench.single(hip, 2048, 2048, 64, 40.0, 0xB1D2)
    p = inst.params
    print("[c2] FP32 2048x2048x64, 300 iterations", flush=True)
    S32 = hip.bld_mult(inst.X.astype(np.float32), 64, p, 300, True,
                       S0=inst.S0.astype(np.float32), dtype=np.float32)
    lm32 = hip.log_marginal_S(S32, p)
    print("[c2] FP64 form", flush=True)
    S64 = hip.bld_mult(inst.X, 64, p, 300, True, S0=inst.S0.astype(np.float32).astype(np.float64))
    lm64 = hip.log_marginal_S(S64, p)
    err = _rel(S32, S64)
    print(f"[c2] rel err {err:.3e}, log-marginal gap {abs(lm32 - lm64) / abs(lm64):.3e}", flush=True)
    assert err <= 1e-4
    assert abs(lm32 - lm64) <= 1e-5 * abs(lm64)


def test_headline_instances_against_oracle(hip, oracle):
    import bench
    c3 = bench.make_c3(hip)
    c3f = [bench.as_dtype(i, hip, np.float32) for i in c3]
    with hip.Batch(c3f, dtype=np.float32, use_psi_appr=True) as b:
        b.solve(200)
        S32 = [b.S(n) for n in range(len(c3))]
    with hip.Batch(c3, dtype=np.float64, use_psi_appr=False) as b:
        b.solve(200)
        S64 = [b.S(n) for n in range(len(c3))]
    gaps = np.array([_rel(S32[n], S64[n]) for n in range(len(c3))])
    worst = [int(n) for n in np.argsort(gaps)[-4:]]
    by_z = {}
    for n, inst in enumerate(c3):
        by_z.setdefault(inst.z, n)
    spread = [by_z[z] for z in sorted(by_z)[::7] if by_z[z] not in worst][:4]
    picks = worst + spread
    assert len(picks) == 8
    for n in picks:
        inst = c3[n]
        print(f"[c3] instance {n}: {inst.X.shape[0]}x{inst.X.shape[1]}x{inst.z}", flush=True)
        ref = oracle.bld_mult(inst.X, inst.z, inst.params.alpha, inst.params.beta, inst.S0,
                              max_iter=200)
        assert _rel(S64[n], ref) <= 1e-10, n
        # the FP32 start is S0 rounded to float; the oracle starts from S0
        assert _rel(S32[n], ref) <= 1e-4, n


def test_serving_workload_requests_equal_alone_and_fp64(hip):
    import bench
    pool = bench.make_pool(hip, 48)
    iters = 200
    with hip.Server(96 << 30, dtype=np.float32, use_psi_appr=True) as srv:
        tickets = [srv.submit(inst, iters) for inst in pool]
        served = []
        for t in tickets:
            assert srv.wait(t, 120.0) == hip.OK
            served.append(srv.collect(t)[0])
    worst = 0.0
    for n, inst in enumerate(pool):
        alone = hip.bld_mult(inst.X, inst.z, inst.params, iters, True, seed=inst.seed,
                             dtype=np.float32)
        assert np.array_equal(alone, served[n]), (n, inst.X.shape, inst.z)
        S64 = hip.bld_mult(inst.X.astype(np.float64), inst.z, inst.params, iters, True,
                           seed=inst.seed, dtype=np.float64)
        err = _rel(served[n], S64)
        worst = max(worst, err)
        assert err <= 1e-4, (n, inst.X.shape, inst.z, err)
        served[n] = None
    print(f"[c4] 48 requests: worst rel err vs FP64 {worst:.3e}", flush=True)
