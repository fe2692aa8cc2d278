"""CPU-side checks of libbnmf_hip.so: it loads, exports every entry point that
include/bnmf_hip.h declares, rejects bad parameters before touching a device
(check_bld_params semantics, src/bld/util_details.hpp:20-28 and
test/bld/bld_mult/test_bld_mult.cpp:10-31), and its host-side input
synthesis follows the allocation model's laws. No kernel is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "bnmf_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bnmf_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared_symbols()
    for n in ("bnmf_bld_mult_f64", "bnmf_bld_mult_f32", "bnmf_batch_create",
              "bnmf_log_marginal_S", "bnmf_tensor_sums"):
        assert n in names


def test_library_exports_every_declared_symbol(hip):
    lib = ctypes.CDLL(hip.LIB_PATH)
    missing = [n for n in _declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version(hip):
    assert hip.lib().bnmf_abi_version() == 100


# ------------------------------------------------------------ parameter checks
def test_bld_mult_parameter_checks(hip):
    # test_bld_mult.cpp:10-31: negative X, wrong z, wrong alpha length throw
    x, y, z = 10, 5, 2
    p = hip.Params.default((x, y, z))
    X = np.full((x, y), 5.0)
    X[0, 0] = -5
    with pytest.raises(ValueError):
        hip.bld_mult(X, z, p)
    with pytest.raises(ValueError):
        hip.bld_mult(np.full((x, y), 5.0), z + 1, p)
    with pytest.raises(ValueError):
        hip.bld_mult(np.full((x + 1, y), 5.0), z, p)
    X = np.full((x, y), 5.0)
    X[1, 1] = np.nan
    with pytest.raises(ValueError):
        hip.bld_mult(X, z, p, max_iter=1)


def test_c_abi_reports_invalid_argument(hip):
    lib = hip.lib()
    X = np.array([[1.0, -2.0]])
    out = np.empty((1, 2, 1))
    a = np.ones(1)
    st = lib.bnmf_bld_mult_f64(X.ctypes.data, 1, 2, 1, a.ctypes.data_as(hip._D),
                               a.ctypes.data_as(hip._D), 5, 0, 1e-50, None, 0, out.ctypes.data)
    assert st == hip.ERR_INVALID_ARGUMENT
    assert b"nonnegative" in lib.bnmf_last_error()
    # like check_bld_params, no upper bound on z (z > 64 runs the wide pass,
    # tested on the GPU); z = 0 is an empty instance with nothing to compute
    good = np.array([[1.0, 2.0]])
    st = lib.bnmf_bld_mult_f64(good.ctypes.data, 1, 2, 0, a.ctypes.data_as(hip._D), None, 5, 0,
                               1e-50, None, 0, None)
    assert st == hip.OK
    st = lib.bnmf_bld_mult_f64(good.ctypes.data, 1, 2, 1 << 31, a.ctypes.data_as(hip._D),
                               a.ctypes.data_as(hip._D), 5, 0, 1e-50, None, 0, out.ctypes.data)
    assert st == hip.ERR_INVALID_ARGUMENT and b"below 2^31" in lib.bnmf_last_error()


def test_sibling_solvers_report_invalid_argument(hip):
    # bld_add and bld_appr check their parameters as bld_mult does
    # (check_bld_params, src/bld/util_details.hpp:20-28) before any device work
    lib = hip.lib()
    a = np.ones(1)
    A = a.ctypes.data_as(hip._D)
    out = np.empty(8)
    O = out.ctypes.data
    good = np.array([[1.0, 2.0]])
    neg = np.array([[1.0, -2.0]])
    calls = {
        "add": lambda X, z: lib.bnmf_bld_add_f64(X.ctypes.data, 1, 2, z, A, A, 5, 1e-50, None, 0, O),
        "appr": lambda X, z: lib.bnmf_bld_appr_f64(X.ctypes.data, 1, 2, z, A, A, 5, 1e-50, None, None,
                                                   0, O, O, O),
    }
    for name, call in calls.items():
        assert call(neg, 1) == hip.ERR_INVALID_ARGUMENT, name
        assert b"nonnegative" in lib.bnmf_last_error(), name
        # z = 0: an empty instance
        assert call(good, 0) == hip.OK, name
        assert call(good, 1 << 31) == hip.ERR_INVALID_ARGUMENT, name
    # null X with a nonempty shape
    st = lib.bnmf_bld_add_f64(None, 1, 2, 1, A, A, 5, 1e-50, None, 0, O)
    assert st == hip.ERR_INVALID_ARGUMENT and b"X is null" in lib.bnmf_last_error()
    st = lib.bnmf_bld_appr_f64(None, 1, 2, 1, A, A, 5, 1e-50, None, None, 0, O, O, O)
    assert st == hip.ERR_INVALID_ARGUMENT and b"X is null" in lib.bnmf_last_error()


def test_log_marginal_parameter_checks(hip):
    # test_alloc_model_funcs.cpp:191-209
    S = np.ones((5, 4, 8))
    p = hip.Params.default((5, 4, 8))
    with pytest.raises(ValueError):
        hip.log_marginal_S(S, hip.Params(p.a, p.b, [1.0] * 6, p.beta))
    with pytest.raises(ValueError):
        hip.log_marginal_S(S, hip.Params(p.a, p.b, p.alpha, [1.0] * 9))


# ------------------------------------------------------------ input synthesis
def test_init_S_is_dirichlet_scaled(hip):
    rng = np.random.default_rng(0)
    X = rng.integers(0, 20, (40, 50)).astype(float)
    S = hip.init_S(7, X, 6)
    assert S.shape == (40, 50, 6)
    assert np.allclose(S.sum(2), X, rtol=1e-13, atol=0)
    assert np.all(S[X == 0] == 0) and np.all(S >= 0)
    frac = S[X > 0] / X[X > 0][:, None]
    assert np.allclose(frac.mean(0), 1 / 6, atol=0.01)  # Dirichlet(1_z) mean
    # deterministic per seed, independent of call order
    assert np.array_equal(S, hip.init_S(7, X, 6))
    assert not np.array_equal(S, hip.init_S(8, X, 6))
    S32 = hip.init_S(7, X.astype(np.float32), 6, dtype=np.float32)
    assert np.allclose(S32, S, rtol=1e-6)


def test_sample_X_moments(hip):
    x, y, z = 60, 400, 5
    p = hip.Params(40.0, 1.0, [1.0] * x, [1.0] * z)
    X = hip.sample_X(3, (x, y, z), p)
    assert np.all(X >= 0) and np.all(X == np.round(X))
    # E[X_ij] = E[L_j] * E[sum_k W_ik H_kj] = a*b / x
    assert X.mean() == pytest.approx(40.0 / x, rel=0.1)
    # column totals X_+j ~ Poisson(L_j) with L_j ~ Gamma(40, 1): mean 40, var 80
    col = X.sum(0)
    assert col.mean() == pytest.approx(40.0, rel=0.05)
    assert col.var() == pytest.approx(80.0, rel=0.3)
    assert np.array_equal(X, hip.sample_X(3, (x, y, z), p))


def test_sample_S_sums_like_sample_X_in_law(hip):
    x, y, z = 30, 300, 4
    p = hip.Params(20.0, 1.0, [1.0] * x, [1.0] * z)
    S = hip.sample_S(1, (x, y, z), p)
    X = hip.sample_X(2, (x, y, z), p)
    assert S.sum(2).mean() == pytest.approx(X.mean(), rel=0.1)
    assert np.all(S == np.round(S))


def test_poisson_large_mean_branch(hip):
    # concentrated model: L ~ Gamma(4000, 1) pushes means far above 12
    x, y, z = 2, 2000, 1
    p = hip.Params(4000.0, 1.0, [1000.0] * x, [1.0] * z)
    X = hip.sample_X(9, (x, y, z), p)
    assert X.mean() == pytest.approx(2000.0, rel=0.02)
    # W_0 ~ 1/2 is shared by all j: Var = W E[L] + W^2 Var(L) = 2000 + 1000
    assert X[0].var() == pytest.approx(3000.0, rel=0.15)
