"""ctypes bindings of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the measured or
shipped path. The C++ source is oracle/bld_oracle.cpp, which restates the
reference CPU solver of eozd/bnmf-algs step by step (file:line citations
there).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_SZ = ctypes.c_size_t


def build(force: bool = False) -> str:
    """Compiles liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for name in ("oracle_psi", "oracle_psi_appr", "oracle_lngamma"):
            getattr(L, name).restype = ctypes.c_double
            getattr(L, name).argtypes = [ctypes.c_double]
        L.oracle_psi_vec.argtypes = [_D, _D, _SZ, ctypes.c_int]
        L.oracle_tensor_sums.argtypes = [_D, _SZ, _SZ, _SZ, _D, _D, _D]
        L.oracle_bld_mult.restype = ctypes.c_int
        L.oracle_bld_mult.argtypes = [_D, _SZ, _SZ, _SZ, _D, _D, _SZ, ctypes.c_int,
                                      ctypes.c_double, _D, _D, ctypes.c_int]
        L.oracle_bld_add.restype = ctypes.c_int
        L.oracle_bld_add.argtypes = [_D, _SZ, _SZ, _SZ, _D, _D, _SZ, ctypes.c_double, _D, _D,
                                     ctypes.c_int]
        L.oracle_bld_add_step.restype = ctypes.c_int
        L.oracle_bld_add_step.argtypes = [_D, _SZ, _SZ, _SZ, _D, _D, ctypes.c_double, _SZ, _D, _D,
                                          _D]
        L.oracle_bld_appr.restype = ctypes.c_int
        L.oracle_bld_appr.argtypes = [_D, _SZ, _SZ, _SZ, _D, _D, _SZ, ctypes.c_double, _D, _D, _D,
                                      _D, _D]
        L.oracle_bld_appr_half.restype = ctypes.c_int
        L.oracle_bld_appr_half.argtypes = [_D, _SZ, _SZ, _SZ, _D, _D, ctypes.c_double, ctypes.c_int,
                                           _D, _D, _D, _D, _D]
        L.oracle_bld_mult_step.restype = ctypes.c_int
        L.oracle_bld_mult_step.argtypes = [_D, _SZ, _SZ, _SZ, _D, _D, ctypes.c_int, ctypes.c_double,
                                           _D, _D, _D, _D, _D, _D, _D, _D, _D]
        L.oracle_log_marginal_S.restype = ctypes.c_double
        L.oracle_log_marginal_S.argtypes = [_D, _SZ, _SZ, _SZ, _D, _D, ctypes.c_double,
                                            ctypes.c_double]
        L.oracle_sparseness.restype = ctypes.c_double
        L.oracle_sparseness.argtypes = [_D, _SZ, _SZ, _SZ]
        L.oracle_bld_fact.argtypes = [_D, _SZ, _SZ, _SZ, _D, _D, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, _D, _D, _D]
        L.oracle_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_D)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def psi(x: float) -> float:
    return lib().oracle_psi(float(x))


def psi_appr(x: float) -> float:
    return lib().oracle_psi_appr(float(x))


def psi_vec(x, appr: bool = False) -> np.ndarray:
    x = _f64(x).ravel()
    out = np.empty_like(x)
    lib().oracle_psi_vec(_p(x), _p(out), x.size, int(appr))
    return out


def tensor_sums(S):
    S = _f64(S)
    x, y, z = S.shape
    pjk, ipk, ijp = np.empty((y, z)), np.empty((x, z)), np.empty((x, y))
    lib().oracle_tensor_sums(_p(S), x, y, z, _p(pjk), _p(ipk), _p(ijp))
    return pjk, ipk, ijp


def bld_mult(X, z, alpha, beta, S0, max_iter=1000, use_psi_appr=False, eps=1e-50, nthreads=0):
    X, alpha, beta, S0 = _f64(X), _f64(alpha), _f64(beta), _f64(S0)
    x, y = X.shape
    assert S0.shape == (x, y, z) and alpha.shape == (x,) and beta.shape == (z,)
    S = np.empty((x, y, z))
    rc = lib().oracle_bld_mult(_p(X), x, y, z, _p(alpha), _p(beta), int(max_iter),
                               int(use_psi_appr), float(eps), _p(S0), _p(S), int(nthreads))
    if rc != 0:
        raise ValueError("oracle_bld_mult: invalid parameters")
    return S


def bld_mult_step(X, alpha, beta, S_in, use_psi_appr=False, eps=1e-50):
    """One iteration; returns dict of S_out and the reference's intermediates."""
    X, alpha, beta, S_in = _f64(X), _f64(alpha), _f64(beta), _f64(S_in)
    x, y, z = S_in.shape
    out = {
        "S": np.empty((x, y, z)), "S_pjk": np.empty((y, z)), "S_ipk": np.empty((x, z)),
        "S_ijp": np.empty((x, y)), "grad_plus": np.empty((x, y, z)),
        "grad_minus": np.empty((x, z)), "nom": np.empty((x, y)), "denom": np.empty((x, y)),
    }
    rc = lib().oracle_bld_mult_step(_p(X), x, y, z, _p(alpha), _p(beta), int(use_psi_appr),
                                    float(eps), _p(S_in), _p(out["S"]), _p(out["S_pjk"]),
                                    _p(out["S_ipk"]), _p(out["S_ijp"]), _p(out["grad_plus"]),
                                    _p(out["grad_minus"]), _p(out["nom"]), _p(out["denom"]))
    if rc != 0:
        raise ValueError("oracle_bld_mult_step: invalid parameters")
    return out


def bld_add(X, z, alpha, beta, S0, max_iter=1000, eps=1e-50, nthreads=0):
    X, alpha, beta, S0 = _f64(X), _f64(alpha), _f64(beta), _f64(S0)
    x, y = X.shape
    assert S0.shape == (x, y, z) and alpha.shape == (x,) and beta.shape == (z,)
    S = np.empty((x, y, z))
    rc = lib().oracle_bld_add(_p(X), x, y, z, _p(alpha), _p(beta), int(max_iter), float(eps),
                              _p(S0), _p(S), int(nthreads))
    if rc != 0:
        raise ValueError("oracle_bld_add: invalid parameters")
    return S


def bld_add_step(X, alpha, beta, S_in, eph=0, eps=1e-50):
    """One bld_add iteration at iteration number eph; returns (S_out, grad_S)."""
    X, alpha, beta, S_in = _f64(X), _f64(alpha), _f64(beta), _f64(S_in)
    x, y, z = S_in.shape
    S, G = np.empty((x, y, z)), np.empty((x, y, z))
    rc = lib().oracle_bld_add_step(_p(X), x, y, z, _p(alpha), _p(beta), float(eps), int(eph),
                                   _p(S_in), _p(S), _p(G))
    if rc != 0:
        raise ValueError("oracle_bld_add_step: invalid parameters")
    return S, G


def bld_appr(X, z, alpha, beta, nu0, mu0, max_iter=1000, eps=1e-50):
    """bld::bld_appr from nu0 (x, z) and mu0 (z, y); returns (S, nu, mu)."""
    X, alpha, beta, nu0, mu0 = _f64(X), _f64(alpha), _f64(beta), _f64(nu0), _f64(mu0)
    x, y = X.shape
    assert nu0.shape == (x, z) and mu0.shape == (z, y)
    S, nu, mu = np.empty((x, y, z)), np.empty((x, z)), np.empty((z, y))
    rc = lib().oracle_bld_appr(_p(X), x, y, z, _p(alpha), _p(beta), int(max_iter), float(eps),
                               _p(nu0), _p(mu0), _p(S), _p(nu), _p(mu))
    if rc != 0:
        raise ValueError("oracle_bld_appr: invalid parameters")
    return S, nu, mu


def bld_appr_half(X, alpha, beta, nu, mu, half, eps=1e-50):
    """One half-step of bld_appr (half 0: nu, 1: mu); returns the updated
    factor and that update's (nom, den)."""
    X, alpha, beta, nu, mu = _f64(X), _f64(alpha), _f64(beta), _f64(nu), _f64(mu)
    x, y = X.shape
    z = nu.shape[1]
    shape = (x, z) if half == 0 else (z, y)
    out, nom, den = np.empty(shape), np.empty(shape), np.empty(shape)
    rc = lib().oracle_bld_appr_half(_p(X), x, y, z, _p(alpha), _p(beta), float(eps), int(half),
                                    _p(nu), _p(mu), _p(out), _p(nom), _p(den))
    if rc != 0:
        raise ValueError("oracle_bld_appr_half: invalid parameters")
    return out, nom, den


def log_marginal_S(S, alpha, beta, a, b) -> float:
    S, alpha, beta = _f64(S), _f64(alpha), _f64(beta)
    x, y, z = S.shape
    return lib().oracle_log_marginal_S(_p(S), x, y, z, _p(alpha), _p(beta), float(a), float(b))


def sparseness(S) -> float:
    S = _f64(S)
    x, y, z = S.shape
    return lib().oracle_sparseness(_p(S), x, y, z)


def bld_fact(S, alpha, beta, a, b, eps=1e-50):
    S, alpha, beta = _f64(S), _f64(alpha), _f64(beta)
    x, y, z = S.shape
    W, H, L = np.empty((x, z)), np.empty((z, y)), np.empty(y)
    lib().oracle_bld_fact(_p(S), x, y, z, _p(alpha), _p(beta), float(a), float(b), float(eps),
                          _p(W), _p(H), _p(L))
    return W, H, L


def num_threads() -> int:
    return lib().oracle_num_threads()
