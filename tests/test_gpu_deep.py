"""Fibers deeper than 64 (the wide pass: one wavefront per fiber, lanes over
k). The reference bounds z nowhere (check_bld_params,
src/bld/util_details.hpp:20-28), so a caller with n_components = 100 must get
the same solver: bld_mult (FP64 and FP32), bld_add, log_marginal_S,
tensor_sums and bld_fact against the FP64 oracle at z in {65, 100, 128, 257},
bld_appr at z in {65, 100, 130}, deep and shallow instances batched together
equal to each solved alone, and a deep request through the server.

Tolerances: FP64 forms 1e-10 relative (normwise) on S and the log-marginal
(bld_appr after one iteration 1e-12, as for shallow fibers); FP32 bulk 1e-4
of the FP64 oracle (bld_appr 5e-5); sums and scoring 1e-12.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEEP = [(40, 70, 65, 30), (33, 50, 100, 40), (20, 37, 128, 25), (12, 19, 257, 20)]


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def _problem(hip, x, y, z, seed, a=40.0):
    p = hip.Params(a, 1.0, [1.0] * x, [1.0] * z)
    X = hip.sample_X(seed, (x, y, z), p)
    S0 = hip.init_S(seed + 1, X, z)
    return p, X, S0


@pytest.mark.parametrize("x,y,z,iters", DEEP)
def test_fp64_deep_matches_oracle(hip, oracle, x, y, z, iters):
    p, X, S0 = _problem(hip, x, y, z, seed=z * 7 + x)
    print(f"[deep] FP64 {x}x{y}x{z}", flush=True)
    for appr in (False, True):
        S_gpu = hip.bld_mult(X, z, p, iters, appr, S0=S0)
        S_cpu = oracle.bld_mult(X, z, p.alpha, p.beta, S0, max_iter=iters, use_psi_appr=appr)
        assert _rel(S_gpu, S_cpu) <= 1e-10, appr
    lm_gpu = hip.log_marginal_S(S_gpu, p)
    lm_cpu = oracle.log_marginal_S(S_cpu, p.alpha, p.beta, p.a, p.b)
    assert abs(lm_gpu - lm_cpu) <= 1e-10 * abs(lm_cpu)


# FP32 against the FP64 oracle: instances with a few rows per component
# amplify FP32 rounding well past 1e-4 (12 x 19 x 257 reaches 1.2e-4), as
# for shallow fibers; these shapes keep x >= 40
DEEP32 = [(40, 70, 65, 30), (41, 50, 100, 40), (48, 37, 128, 25), (60, 29, 257, 20)]


@pytest.mark.parametrize("x,y,z,iters", DEEP32)
def test_fp32_deep_within_1e4_of_oracle(hip, oracle, x, y, z, iters):
    p, X, S0 = _problem(hip, x, y, z, seed=z * 5 + y)
    S32 = hip.bld_mult(X.astype(np.float32), z, p, iters, True, S0=S0.astype(np.float32),
                       dtype=np.float32)
    S64 = oracle.bld_mult(X, z, p.alpha, p.beta, S0.astype(np.float32).astype(np.float64),
                          max_iter=iters, use_psi_appr=True)
    assert _rel(S32, S64) <= 1e-4
    lm32 = hip.log_marginal_S(S32, p)
    lm64 = oracle.log_marginal_S(S64, p.alpha, p.beta, p.a, p.b)
    assert abs(lm32 - lm64) <= 1e-4 * abs(lm64)


def test_fp64_deep_nonuniform_parameters_and_zero_fibers(hip, oracle):
    # beta < 1 takes psi(S+1)'s general branch; zero rows of X give zero fibers
    x, y, z = 30, 41, 90
    rng = np.random.default_rng(3)
    p = hip.Params(5.0, 2.0, rng.uniform(0.3, 2.0, x).tolist(), rng.uniform(0.5, 3.0, z).tolist())
    X = hip.sample_X(4, (x, y, z), p)
    X[3, :] = 0.0
    X[:, 7] = 0.0
    S0 = hip.init_S(5, X, z)
    S_gpu = hip.bld_mult(X, z, p, 25, False, S0=S0)
    S_cpu = oracle.bld_mult(X, z, p.alpha, p.beta, S0, max_iter=25)
    assert _rel(S_gpu, S_cpu) <= 1e-10
    assert not S_gpu[3].any() and not S_gpu[:, 7].any()
    S32 = hip.bld_mult(X.astype(np.float32), z, p, 25, True, S0=S0.astype(np.float32), dtype=np.float32)
    assert _rel(S32, S_cpu) <= 1e-4


@pytest.mark.parametrize("z", [65, 130])
def test_bld_add_deep_matches_oracle(hip, oracle, z):
    x, y = 26, 45
    p, X, S0 = _problem(hip, x, y, z, seed=z)
    S_gpu = hip.bld_add(X, z, p, 40, S0=S0)
    S_cpu = oracle.bld_add(X, z, p.alpha, p.beta, S0, max_iter=40)
    assert _rel(S_gpu, S_cpu) <= 1e-10
    S32 = hip.bld_add(X.astype(np.float32), z, p, 40, S0=S0.astype(np.float32), dtype=np.float32)
    S64 = oracle.bld_add(X, z, p.alpha, p.beta, S0.astype(np.float32).astype(np.float64), max_iter=40)
    assert _rel(S32, S64) <= 1e-5


@pytest.mark.parametrize("z", [65, 100, 257])
def test_log_marginal_sums_and_fact_deep(hip, oracle, z):
    x, y = 21, 34
    rng = np.random.default_rng(z)
    S = rng.gamma(0.5, 2.0, (x, y, z))
    p = hip.Params(7.0, 2.0, list(rng.uniform(0.5, 3, x)), list(rng.uniform(1.0, 3, z)))
    ref = oracle.log_marginal_S(S, p.alpha, p.beta, p.a, p.b)
    assert hip.log_marginal_S(S, p) == pytest.approx(ref, rel=1e-12)
    S32 = S.astype(np.float32)
    ref32 = oracle.log_marginal_S(S32.astype(np.float64), p.alpha, p.beta, p.a, p.b)
    assert hip.log_marginal_S(S32, p) == pytest.approx(ref32, rel=1e-12)
    pjk, ipk, ijp = hip.tensor_sums(S)
    rp, ri, rij = oracle.tensor_sums(S)
    assert _rel(pjk, rp) < 1e-13 and _rel(ipk, ri) < 1e-13 and _rel(ijp, rij) < 1e-14
    W, H, L = hip.bld_fact(S, p)
    Wr, Hr, Lr = oracle.bld_fact(S, p.alpha, p.beta, p.a, p.b, 1e-50)
    for got, want in ((W, Wr), (H, Hr), (L, Lr)):
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_deep_and_shallow_batched_equal_alone(hip, dtype):
    shapes = [(40, 33, 100), (30, 70, 17), (25, 65, 65), (17, 40, 8)]
    insts = []
    for n, (x, y, z) in enumerate(shapes):
        p, X, S0 = _problem(hip, x, y, z, seed=300 + n)
        insts.append(hip.Instance(X=X.astype(dtype), z=z, params=p, S0=S0.astype(dtype)))
    appr = dtype == np.float32
    with hip.Batch(insts, dtype=dtype, use_psi_appr=appr) as b:
        b.solve(19)
        mixed = [b.S(n) for n in range(len(insts))]
        lms = b.log_marginal()
    for n, inst in enumerate(insts):
        alone = hip.bld_mult(inst.X, inst.z, inst.params, 19, appr, S0=inst.S0, dtype=dtype)
        assert np.array_equal(alone, mixed[n]), shapes[n]
        assert lms[n] == pytest.approx(hip.log_marginal_S(alone, inst.params), rel=1e-13)


def test_deep_request_served_equals_alone(hip):
    p, X, _ = _problem(hip, 50, 60, 96, seed=77)
    inst = hip.Instance(X=X.astype(np.float32), z=96, params=p, seed=78)
    with hip.Server(1 << 30, dtype=np.float32, use_psi_appr=True) as srv:
        t = srv.submit(inst, 30)
        assert srv.wait(t, 60.0) == hip.OK
        served = srv.collect(t)[0]
    alone = hip.bld_mult(inst.X, 96, p, 30, True, seed=78, dtype=np.float32)
    assert np.array_equal(served, alone)


def test_empty_depth(hip):
    # z = 0 (an empty beta) is an empty instance, as x = 0 or y = 0 are
    X = np.ones((3, 4))
    p = hip.Params(1.0, 1.0, [1.0] * 3, [])
    S = hip.bld_mult(X, 0, p, 5)
    assert S.shape == (3, 4, 0)


class _ApprInst:
    def __init__(self, X, z, params, nu0, mu0):
        self.X, self.z, self.params, self.nu0, self.mu0, self.seed = X, z, params, nu0, mu0, 0


def _appr_problem(hip, x, y, z, seed):
    p = hip.Params(40.0, 1.0, [1.0] * x, [1.0] * z)
    X = hip.sample_X(seed, (x, y, z), p)
    rng = np.random.default_rng(seed)
    nu0 = rng.dirichlet(np.ones(z), size=x)
    mu0 = rng.dirichlet(np.ones(z), size=y).T.copy()
    return p, X, nu0, mu0


APPR_DEEP = [(40, 70, 65), (33, 50, 100), (20, 37, 130)]


@pytest.mark.parametrize("x,y,z", APPR_DEEP)
def test_appr_fp64_deep_matches_oracle(hip, oracle, x, y, z):
    p, X, nu0, mu0 = _appr_problem(hip, x, y, z, z + 1)
    print(f"[deep] bld_appr FP64 {x}x{y}x{z}", flush=True)
    S, nu, mu = hip.bld_appr(X, z, p, 1, nu0=nu0, mu0=mu0)
    Sr, nur, mur = oracle.bld_appr(X, z, p.alpha, p.beta, nu0, mu0, max_iter=1)
    assert _rel(nu, nur) < 1e-12 and _rel(mu, mur) < 1e-12 and _rel(S, Sr) < 1e-12
    S, nu, mu = hip.bld_appr(X, z, p, 30, nu0=nu0, mu0=mu0)
    Sr, nur, mur = oracle.bld_appr(X, z, p.alpha, p.beta, nu0, mu0, max_iter=30)
    assert _rel(nu, nur) < 1e-10 and _rel(mu, mur) < 1e-10 and _rel(S, Sr) < 1e-10
    # rows of nu and columns of mu stay on the simplex, S_ij+ = X
    assert np.abs(nu.sum(1) - 1).max() < 1e-12 and np.abs(mu.sum(0) - 1).max() < 1e-12
    assert np.abs(S.sum(2) - X).max() <= 1e-12 * max(X.max(), 1)


def test_appr_fp32_deep_close_to_fp64_oracle(hip, oracle):
    x, y, z = 41, 50, 100
    p, X, nu0, mu0 = _appr_problem(hip, x, y, z, 9)
    nu0f, mu0f = nu0.astype(np.float32), mu0.astype(np.float32)
    S, _, _ = hip.bld_appr(X.astype(np.float32), z, p, 60, nu0=nu0f, mu0=mu0f, dtype=np.float32)
    Sr, _, _ = oracle.bld_appr(X, z, p.alpha, p.beta, nu0f.astype(np.float64),
                               mu0f.astype(np.float64), max_iter=60)
    assert _rel(S, Sr) < 5e-5, _rel(S, Sr)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_appr_deep_and_shallow_batched_equal_alone(hip, dtype):
    shapes = [(30, 45, 100), (25, 70, 17), (20, 33, 65)]
    insts, alone = [], []
    for n, (x, y, z) in enumerate(shapes):
        p, X, nu0, mu0 = _appr_problem(hip, x, y, z, 50 + n)
        X, nu0, mu0 = X.astype(dtype), nu0.astype(dtype), mu0.astype(dtype)
        insts.append(_ApprInst(X, z, p, nu0, mu0))
        alone.append(hip.bld_appr(X, z, p, 10, nu0=nu0, mu0=mu0, dtype=dtype))
    with hip.ApprBatch(insts, dtype=dtype) as b:
        b.solve(10)
        for n in range(len(shapes)):
            for g, a in zip(b.get(n), alone[n]):
                assert np.array_equal(g, a), shapes[n]
