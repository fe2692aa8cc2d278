"""GPU parity of the HIP solver against the FP64 oracle and golden vectors.

Tolerances (written per test):
  * plain FP64 form vs the oracle: 1e-10 relative (normwise, as Eigen's
    isApprox) on S and on log_marginal_S;
  * FP32-bulk form vs the FP64 oracle: 1e-4 relative normwise on S;
  * device digamma forms vs mpmath: see each test;
  * batched vs single-instance runs: bitwise equal.
Inputs are seeded allocation-model samples (bnmf_hip.sample_X) with seeded
init_S tensors, passed identically to the GPU and to the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)


def _elem_rel(a, b, floor=1e-6):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    den = np.maximum(np.abs(b), floor * max(np.abs(b).max(), 1e-300))
    return float((np.abs(a - b) / den).max())


def _problem(hip, x, y, z, seed, a=40.0, dtype=np.float64):
    p = hip.Params(a, 1.0, [1.0] * x, [1.0] * z)
    X = hip.sample_X(seed, (x, y, z), p)
    S0 = hip.init_S(seed + 1, X, z)
    return p, X.astype(dtype), S0.astype(dtype)


# ------------------------------------------------------------------ digamma
def test_device_psi_exact(hip, golden):
    xs = np.array(golden["psi"]["x"])
    ref = np.array(golden["psi"]["psi"])
    got = hip.psi_eval(1, xs)
    assert np.all(np.abs(got - ref) <= 3e-14 * np.maximum(1, np.abs(ref)))
    neg = hip.psi_eval(1, golden["psi_neg"]["x"])
    assert np.allclose(neg, golden["psi_neg"]["psi"], rtol=1e-12, atol=1e-12)


def test_device_psi_sp1_f64(hip, golden):
    # the FP64 update's psi(S+1): a rational fit on [1, 2) reached by
    # recurrence from x < 10, psi_f64 elsewhere; against mpmath on a dense
    # grid (every shift count, both ends of each unit interval) and the
    # golden vectors at x >= 1
    import mpmath
    xs = np.concatenate([np.linspace(1.0, 10.0, 1801)[:-1], 1.0 + np.arange(1, 10) * 1e-12,
                         np.arange(2.0, 10.0) - 1e-13, [1.4616321449683622, 9.999999999, 10.0, 37.5, 1e6]])
    ref = np.array([float(mpmath.digamma(mpmath.mpf(float(v)))) for v in xs])
    got = hip.psi_eval(4, xs)
    assert np.all(np.abs(got - ref) <= 8e-16 * np.maximum(1.0, np.abs(ref)))
    gx = np.array(golden["psi"]["x"])
    gp = np.array(golden["psi"]["psi"])
    keep = gx >= 1.0
    got = hip.psi_eval(4, gx[keep])
    assert np.all(np.abs(got - gp[keep]) <= 3e-14 * np.maximum(1, np.abs(gp[keep])))


def test_device_psi_appr(hip, golden):
    xs = np.array(golden["psi_appr"]["x"])
    ref = np.array(golden["psi_appr"]["psi_appr"])
    got = hip.psi_eval(2, xs)
    assert np.all(np.abs(got - ref) <= 4e-15 * np.maximum(1, np.abs(ref)))


def test_device_psi_fp32(hip, oracle, golden):
    xs = np.array(golden["psi"]["x"])
    xs = xs[(xs > 1e-3) & (xs < 1e7)]
    # reference at the float32-rounded argument (the oracle's psi is pinned
    # to mpmath to 4e-15)
    xr = xs.astype(np.float32).astype(np.float64)
    ref = oracle.psi_vec(xr)
    got = hip.psi_eval(0, xs)
    assert np.all(np.abs(got - ref) <= 6e-7 * np.maximum(1, np.abs(ref)))
    hot = xr[xr >= 1]
    ref_hot = oracle.psi_vec(hot)
    for kind in (3,):
        got_hot = hip.psi_eval(kind, hot)
        assert np.all(np.abs(got_hot - ref_hot) <= 6e-7 * np.maximum(1, np.abs(ref_hot))), kind


# ------------------------------------------------------------------ FP64
@pytest.mark.parametrize("mode", ["exact", "appr"])
@pytest.mark.parametrize("case", [0, 1])
def test_fp64_tiny_iterations_match_mpmath(hip, golden, mode, case):
    c = golden["iterations"][case]
    X, S0 = np.array(c["X"]), np.array(c["S0"])
    p = hip.Params(c["a"], c["b"], c["alpha"], c["beta"])
    z = len(c["beta"])
    S1 = hip.bld_mult(X, z, p, 1, mode == "appr", S0=S0)
    assert _rel(S1, c[mode]["S1"]) < 1e-13
    S5 = hip.bld_mult(X, z, p, 5, mode == "appr", S0=S0)
    assert _rel(S5, c[mode]["S5"]) < 1e-12


SHAPES_F64 = [
    (100, 100, 17, 1000),  # reference CPU config, full length
    (356, 478, 17, 20),    # test_bld_mult_cuda_funcs.cpp:61
    (1024, 456, 17, 10),   # test/cuda/test_tensor_ops.cpp:17
    (37, 65, 5, 50),
    (65, 129, 33, 20),
    (200, 70, 64, 15),
    (13, 200, 8, 30),
    (1, 1, 1, 3),
    (3, 1, 2, 7),
    (50, 77, 6, 12),       # z = 2 mod 4 with odd y: rows start misaligned
    (33, 99, 10, 12),
    (70, 91, 12, 12),      # z = 4 mod 8
    (40, 71, 24, 10),
    (30, 65, 48, 8),
    (20, 67, 16, 10),
    (9, 130, 2, 20),
    (17, 66, 4, 20),
]


@pytest.mark.parametrize("x,y,z,iters", SHAPES_F64)
def test_fp64_matches_oracle(hip, oracle, x, y, z, iters):
    p, X, S0 = _problem(hip, x, y, z, seed=x * 1000 + y + z)
    S_gpu = hip.bld_mult(X, z, p, iters, False, S0=S0)
    S_cpu = oracle.bld_mult(X, z, p.alpha, p.beta, S0, max_iter=iters)
    assert _rel(S_gpu, S_cpu) <= 1e-10
    lm_gpu = hip.log_marginal_S(S_gpu, p)
    lm_cpu = oracle.log_marginal_S(S_cpu, p.alpha, p.beta, p.a, p.b)
    assert abs(lm_gpu - lm_cpu) <= 1e-10 * abs(lm_cpu)


def test_fp64_psi_appr_matches_oracle(hip, oracle):
    x, y, z, iters = 100, 100, 17, 300
    p, X, S0 = _problem(hip, x, y, z, seed=77)
    S_gpu = hip.bld_mult(X, z, p, iters, True, S0=S0)
    S_cpu = oracle.bld_mult(X, z, p.alpha, p.beta, S0, max_iter=iters, use_psi_appr=True)
    assert _rel(S_gpu, S_cpu) <= 1e-10


def test_fp64_nonuniform_params(hip, oracle):
    # alpha/beta differ per index, so a transposed alpha (x) / beta (z)
    # lookup cannot pass
    rng = np.random.default_rng(8)
    x, y, z = 70, 90, 6
    p = hip.Params(10.0, 2.0, list(rng.uniform(0.05, 3, x)), list(rng.uniform(0.5, 4, z)))
    X = hip.sample_X(5, (x, y, z), p)
    S0 = hip.init_S(6, X, z)
    S_gpu = hip.bld_mult(X, z, p, 40, False, S0=S0)
    S_cpu = oracle.bld_mult(X, z, p.alpha, p.beta, S0, max_iter=40)
    assert _rel(S_gpu, S_cpu) <= 1e-10


@pytest.mark.parametrize("appr", [False, True])
def test_fixture_8x10_gpu(hip, oracle, golden, appr):
    fx = golden["fixture_8x10"]
    X, S0 = np.array(fx["X"]), np.array(fx["S0"])
    p = hip.Params(fx["a"], fx["b"], fx["alpha"], fx["beta"])
    S = hip.bld_mult(X, fx["z"], p, fx["max_iter"], appr, S0=S0)
    th = fx["thresholds"]
    assert hip.log_marginal_S(S, p) >= th["log_marginal_min"]
    assert oracle.sparseness(S) >= th["sparseness_min"]
    sij = S.sum(2)
    assert np.linalg.norm(X - sij) / min(np.linalg.norm(X), np.linalg.norm(sij)) <= (
        th["isapprox_appr"] if appr else th["isapprox_exact"])
    S_cpu = oracle.bld_mult(X, fx["z"], fx["alpha"], fx["beta"], S0, max_iter=fx["max_iter"],
                            use_psi_appr=appr)
    assert _rel(S, S_cpu) <= 1e-10


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_zero_matrix_gives_zero(hip, dtype):
    p = hip.Params.default((8, 10, 3))
    S = hip.bld_mult(np.zeros((8, 10)), 3, p, 50, dtype=dtype)
    assert np.all(S == 0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_z1_fixed_point(hip, dtype):
    rng = np.random.default_rng(1)
    X = rng.integers(0, 9, (66, 70)).astype(dtype)
    p = hip.Params.default((66, 70, 1))
    S = hip.bld_mult(X, 1, p, 30, dtype=dtype, S0=X[:, :, None])
    tol = 1e-13 if dtype == np.float64 else 1e-6
    assert np.allclose(S[:, :, 0], X, rtol=tol, atol=0)


def test_zero_fibers_stay_zero_and_finite(hip):
    p, X, S0 = _problem(hip, 150, 130, 12, seed=3)
    X[5, :] = 0
    X[:, 7] = 0
    S0 = hip.init_S(4, X, 12)
    for dt in (np.float64, np.float32):
        S = hip.bld_mult(X.astype(dt), 12, p, 25, dtype=dt, S0=S0.astype(dt))
        assert np.all(S[X == 0] == 0)
        assert np.all(np.isfinite(S))


# ------------------------------------------------------------------ FP32
@pytest.mark.parametrize("x,y,z,iters", [(256, 256, 32, 100), (300, 200, 17, 100),
                                         (128, 640, 64, 60), (90, 70, 5, 200),
                                         (77, 93, 6, 80), (61, 131, 12, 80),
                                         (45, 77, 24, 60), (39, 101, 48, 40),
                                         (50, 67, 16, 80), (70, 99, 8, 80),
                                         (33, 65, 2, 80), (40, 129, 62, 30)])
def test_fp32_bulk_within_1e4_of_fp64_oracle(hip, oracle, x, y, z, iters):
    p, X, S0 = _problem(hip, x, y, z, seed=11 * x + z)
    S32 = hip.bld_mult(X.astype(np.float32), z, p, iters, True, S0=S0.astype(np.float32),
                       dtype=np.float32)
    S64 = oracle.bld_mult(X, z, p.alpha, p.beta, S0.astype(np.float32).astype(np.float64),
                          max_iter=iters, use_psi_appr=True)
    assert _rel(S32, S64) <= 1e-4
    lm32 = hip.log_marginal_S(S32, p)
    lm64 = oracle.log_marginal_S(S64, p.alpha, p.beta, p.a, p.b)
    assert abs(lm32 - lm64) <= 1e-4 * abs(lm64)


# ------------------------------------------------------------ scoring / sums
@pytest.mark.parametrize("shape", [(64, 64, 8), (70, 130, 17), (5, 300, 64), (1, 1, 1)])
def test_log_marginal_matches_oracle(hip, oracle, shape):
    x, y, z = shape
    rng = np.random.default_rng(x + y + z)
    S = rng.gamma(0.5, 2.0, shape)
    p = hip.Params(7.0, 2.0, list(rng.uniform(0.1, 3, x)), list(rng.uniform(0.1, 3, z)))
    ref = oracle.log_marginal_S(S, p.alpha, p.beta, p.a, p.b)
    assert hip.log_marginal_S(S, p) == pytest.approx(ref, rel=1e-12)
    S32 = S.astype(np.float32)
    ref32 = oracle.log_marginal_S(S32.astype(np.float64), p.alpha, p.beta, p.a, p.b)
    assert hip.log_marginal_S(S32, p) == pytest.approx(ref32, rel=1e-12)


def test_tensor_sums_match_oracle(hip, oracle):
    # test/cuda/test_tensor_ops.cpp:13-69 shape
    rng = np.random.default_rng(0)
    S = rng.uniform(-1, 1, (1024, 456, 17))
    pjk, ipk, ijp = hip.tensor_sums(S)
    rp, ri, rij = oracle.tensor_sums(S)
    assert _rel(pjk, rp) < 1e-13 and _rel(ipk, ri) < 1e-13 and _rel(ijp, rij) < 1e-14
    p32, i32, ij32 = hip.tensor_sums(S.astype(np.float32))
    assert _rel(p32, rp) < 1e-6 and _rel(i32, ri) < 1e-6 and _rel(ij32, rij) < 1e-6


# ------------------------------------------------------- batch / determinism
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_alone_vs_mix_bitwise(hip, dtype):
    shapes = [(40, 33, 4), (130, 70, 17), (64, 64, 32), (7, 300, 9), (257, 65, 64)]
    insts = []
    for n, (x, y, z) in enumerate(shapes):
        p, X, S0 = _problem(hip, x, y, z, seed=100 + n)
        insts.append(hip.Instance(X=X.astype(dtype), z=z, params=p, S0=S0.astype(dtype)))
    iters = 23
    appr = dtype == np.float32
    with hip.Batch(insts, dtype=dtype, use_psi_appr=appr) as b:
        b.solve(iters)
        mixed = [b.S(n) for n in range(len(insts))]
        lms = b.log_marginal()
    for n, inst in enumerate(insts):
        alone = hip.bld_mult(inst.X, inst.z, inst.params, iters, appr, S0=inst.S0, dtype=dtype)
        assert np.array_equal(alone, mixed[n]), shapes[n]
        assert lms[n] == pytest.approx(hip.log_marginal_S(alone, inst.params), rel=1e-14)


def test_repeat_solves_are_bitwise_identical(hip):
    p, X, S0 = _problem(hip, 300, 200, 20, seed=9)
    insts = [hip.Instance(X=X.astype(np.float32), z=20, params=p, S0=S0.astype(np.float32))]
    with hip.Batch(insts, dtype=np.float32) as b:
        b.solve(40)
        first = b.S(0)
        b.solve(40)
        assert np.array_equal(first, b.S(0))


def test_seeded_init_matches_explicit_S0(hip):
    # the seeded start is drawn on the device; it agrees with the host's
    # init_S draw to the last bits of the double log (the two logs may
    # differ by an ulp), so the solves agree to rounding, not bitwise
    p, X, _ = _problem(hip, 50, 60, 7, seed=21)
    S_seed = hip.bld_mult(X, 7, p, 12, seed=1234)
    S_expl = hip.bld_mult(X, 7, p, 12, S0=hip.init_S(1234, X, 7))
    assert np.allclose(S_seed, S_expl, rtol=1e-12, atol=0)


# ------------------------------------------------------------------ bld_fact
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_bld_fact_matches_oracle(hip, oracle, dtype):
    x, y, z = 70, 130, 17
    rng = np.random.default_rng(21)
    p = hip.Params(7.0, 2.0, rng.uniform(0.5, 2.0, x).tolist(), rng.uniform(1.0, 3.0, z).tolist())
    X = hip.sample_X(22, (x, y, z), p)
    inst = hip.Instance(X=X.astype(dtype), z=z, params=p, seed=23)
    with hip.Batch([inst], dtype=dtype, use_psi_appr=True) as b:
        b.solve(30)
        S = b.S(0)
        W, H, L = b.fact(0, 1e-50)
        # again after scoring (the partial sums are then recomputed from S)
        b.log_marginal()
        W2, H2, L2 = b.fact(0, 1e-50)
    Wr, Hr, Lr = oracle.bld_fact(S.astype(np.float64), p.alpha, p.beta, p.a, p.b, 1e-50)
    tol = 1e-12 if dtype == np.float64 else 2e-6
    def close(got, ref):  # normwise: entries can be near zero (alpha + S - 1)
        return np.max(np.abs(got - ref)) <= tol * np.max(np.abs(ref))
    for got, ref in ((W, Wr), (H, Hr), (L, Lr), (W2, Wr), (H2, Hr), (L2, Lr)):
        assert close(got, ref)
    Wo, Ho, Lo = hip.bld_fact(S, p)
    for got, ref in ((Wo, Wr), (Ho, Hr), (Lo, Lr)):
        assert close(got, ref)


def test_fp32_batched_workload_within_1e4_of_fp64(hip):
    """The benchmark's 256-instance workload (x, y ~ U{32..512}, z ~ U{4..32},
    200 iterations): FP32 bulk within 1e-4 normwise of the FP64 form with
    exact digamma on every instance, both solved as batches (the FP64 batch
    equals the FP64 single solves bit for bit)."""
    import bench
    c3 = bench.make_c3(hip)
    c3f = [bench.as_dtype(i, hip, np.float32) for i in c3]
    with hip.Batch(c3f, dtype=np.float32, use_psi_appr=True) as b:
        b.solve(200)
        S32 = [b.S(n) for n in range(len(c3))]
    with hip.Batch(c3, dtype=np.float64, use_psi_appr=False) as b:
        b.solve(200)
        errs = np.array([bench.rel_err(S32[n], b.S(n)) for n in range(len(c3))])
    assert errs.max() <= 1e-4, (errs.max(), int(errs.argmax()))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_draws_S0_on_device_like_host_init(hip, dtype):
    # instances without S0 draw it on the GPU (init_S, the host's Philox
    # stream); a zero-iteration solve returns it
    shapes = [(70, 90, 11), (33, 130, 64), (129, 64, 5)]
    insts = []
    for n, (x, y, z) in enumerate(shapes):
        p = hip.Params(40.0, 1.0, [1.0] * x, [1.0] * z)
        X = hip.sample_X(50 + n, (x, y, z), p).astype(dtype)
        insts.append(hip.Instance(X=X, z=z, params=p, seed=900 + n))
    with hip.Batch(insts, dtype=dtype) as b:
        b.solve(0)
        for n, inst in enumerate(insts):
            got = b.S(n)
            ref = hip.init_S(inst.seed, inst.X.astype(np.float64), inst.z)
            tol = 1e-13 if dtype == np.float64 else 1e-6
            assert np.allclose(got, ref, rtol=tol, atol=0), shapes[n]
            assert np.allclose(got.sum(2), inst.X, rtol=10 * tol, atol=0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_convergence_check(hip, oracle, dtype):
    shapes = [(70, 90, 11), (40, 130, 24), (129, 64, 5)]
    insts = []
    for n, (x, y, z) in enumerate(shapes):
        p, X, S0 = _problem(hip, x, y, z, seed=60 + n, dtype=dtype)
        insts.append(hip.Instance(X=X, z=z, params=p, S0=S0))
    with hip.Batch(insts, dtype=dtype) as b:
        # tol = 0 never stops early, and the checks do not perturb the solve
        b.solve(95)
        ref = [b.S(n) for n in range(len(shapes))]
        lm_ref = b.log_marginal()
        it, lm = b.solve_tol(95, 0.0, check_every=10)
        assert it == 95
        for n in range(len(shapes)):
            assert np.array_equal(b.S(n), ref[n]), shapes[n]
        assert np.array_equal(lm, lm_ref)
        # a loose tolerance stops early, at a multiple of check_every, with
        # the last two checks within tol of each other
        it2, lm2 = b.solve_tol(3000, 1e-5, check_every=20)
        assert it2 < 3000 and it2 % 20 == 0
        b.solve(it2 - 20)
        lm_prev = b.log_marginal()
        assert np.all(np.abs(lm2 - lm_prev) <= 1e-5 * np.abs(lm2))
        # the log-marginals returned are those of the returned S
        b.solve(it2)
        assert np.array_equal(b.log_marginal(), lm2)


@pytest.mark.parametrize("algo", ["mult", "add"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_graph_replays_any_iteration_count(hip, dtype, algo):
    # a solve replays captured graphs of 32, 16, ..., 1 iterations (the count
    # in binary); one-iteration chunks through the convergence check give the
    # same S bit for bit, for counts on both sides of the graph sizes
    shapes = [(70, 90, 11), (40, 130, 24)]
    insts = []
    for n, (x, y, z) in enumerate(shapes):
        p, X, S0 = _problem(hip, x, y, z, seed=80 + n, dtype=dtype)
        insts.append(hip.Instance(X=X, z=z, params=p, S0=S0))
    code = hip.ALGO_ADD if algo == "add" else hip.ALGO_MULT
    with hip.Batch(insts, dtype=dtype, algo=code) as b:
        for iters in (1, 7, 31, 33, 64, 95):
            b.solve(iters)
            ref = [b.S(n) for n in range(len(shapes))]
            it, _ = b.solve_tol(iters, 0.0, check_every=1)
            assert it == iters
            for n in range(len(shapes)):
                assert np.array_equal(b.S(n), ref[n]), (iters, shapes[n])


def test_device_sampler_matches_host_sampler(hip):
    # the Poisson draws on the GPU from the host sampler's streams: equal
    # draws (a last-bit difference of exp / log / lgamma at an acceptance
    # boundary could flip one; none does for these seeds), both regimes of
    # the draw (means below and above 12)
    for seed, (x, y, z), a in [(5, (200, 300, 17), 40.0), (6, (64, 90, 8), 400.0)]:
        p = hip.Params(a, 1.0, [1.0] * x, [1.0] * z)
        X_h = hip.sample_X(seed, (x, y, z), p)
        X_d = hip.sample_X_gpu(seed, (x, y, z), p)
        assert np.array_equal(X_h, X_d)
    p = hip.Params(40.0, 1.0, [1.0] * 50, [1.0] * 9)
    S_h = hip.sample_S(11, (50, 60, 9), p)
    S_d = hip.sample_S_gpu(11, (50, 60, 9), p)
    assert np.array_equal(S_h, S_d)
    assert S_d.min() >= 0 and np.all(S_d == np.round(S_d))


@pytest.mark.parametrize("z", [12, 33, 37, 45, 52, 60])
@pytest.mark.parametrize("x", [1, 3, 5, 9, 131])
def test_fp32_few_rows_and_large_z(hip, oracle, x, z):
    # exact-z FP32 kernels at one and several waves per SIMD with row counts
    # per wave of 0, 1, 2 and many (odd and even), including x = 1, where
    # grad_plus and grad_minus vanish (beta = 1) and the update is 0/0
    # without the double eps guard of the reference. Few iterations: an
    # indexing mistake shows at once, while few-row instances amplify FP32
    # rounding over many iterations
    y = 70
    p, X, S0 = _problem(hip, x, y, z, seed=7 * x + z)
    S = hip.bld_mult(X.astype(np.float32), z, p, 5, True, S0=S0.astype(np.float32),
                     dtype=np.float32)
    ref = oracle.bld_mult(X, z, p.alpha, p.beta, S0.astype(np.float32).astype(np.float64),
                          max_iter=5, use_psi_appr=True)
    assert _rel(S, ref) < 1e-4, _rel(S, ref)
    Sa = hip.bld_add(X.astype(np.float32), z, p, 5, S0=S0.astype(np.float32), dtype=np.float32)
    refa = oracle.bld_add(X, z, p.alpha, p.beta, S0.astype(np.float32).astype(np.float64),
                          max_iter=5)
    assert _rel(Sa, refa) < 1e-4, _rel(Sa, refa)
