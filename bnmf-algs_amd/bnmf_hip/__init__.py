"""Python host mirror of the MI355X-native multiplicative BLD solver.

Thin ctypes layer over lib/libbnmf_hip.so (C ABI: include/bnmf_hip.h). The
names follow the reference's C++ API (eozd/bnmf-algs):

  bld_mult / bld_mult_cuda   bld::bld_mult, bld::bld_mult_cuda
                             (src/bld/bld_mult/bld_mult.hpp:83-86, :226-229)
  Params                     alloc_model::Params
                             (src/alloc_model/alloc_model_params.hpp:25-84)
  log_marginal_S             alloc_model::log_marginal_S
                             (src/alloc_model/alloc_model_funcs.hpp:574-586)
  tensor_sums                cuda::tensor_sums (src/cuda/tensor_ops.hpp:40-42)
  bnmf_priors / sample_S     alloc_model_funcs.hpp:463-552 (seeded)
  init_S                     bld_mult_funcs.hpp:36-74 (seeded)

Parameter errors raise ValueError (the reference throws
std::invalid_argument under BUILD_TESTING, src/defs.hpp:12-16); device
errors raise RuntimeError. Every solver call runs on the GPU; there is no
CPU path, and importing this module fails if the library is not built.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# BNMF_HIP_LIB selects another build of the library (e.g. for comparisons)
LIB_PATH = os.environ.get("BNMF_HIP_LIB") or os.path.join(os.path.dirname(_HERE), "lib",
                                                           "libbnmf_hip.so")

F32, F64 = 0, 1
OK, ERR_INVALID_ARGUMENT, ERR_OUT_OF_MEMORY, ERR_DEVICE = 0, 1, 2, 3
ERR_NOT_READY, ERR_NOT_FOUND, ERR_REJECTED, ERR_TIMEOUT = 4, 5, 6, 7

_D = ctypes.POINTER(ctypes.c_double)
_SZ = ctypes.c_size_t
_VP = ctypes.c_void_p


class bnmf_bld_desc(ctypes.Structure):
    _fields_ = [("x", _SZ), ("y", _SZ), ("z", _SZ), ("X", _VP), ("alpha", _D), ("beta", _D),
                ("a", ctypes.c_double), ("b", ctypes.c_double), ("S0", _VP),
                ("seed", ctypes.c_uint64)]


class bnmf_appr_desc(ctypes.Structure):
    _fields_ = [("x", _SZ), ("y", _SZ), ("z", _SZ), ("X", _VP), ("alpha", _D), ("beta", _D),
                ("nu0", _VP), ("mu0", _VP), ("seed", ctypes.c_uint64)]


class bnmf_batch_info(ctypes.Structure):
    _fields_ = [("n_instances", _SZ), ("elements", _SZ), ("fibers", _SZ),
                ("device_bytes", _SZ), ("pass_items", _SZ), ("prep_items", _SZ),
                ("n_buckets", ctypes.c_int)]


class bnmf_request_stats(ctypes.Structure):
    _fields_ = [("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double),
                ("done_s", ctypes.c_double), ("iterations", _SZ), ("device_bytes", _SZ)]


class bnmf_server_stats(ctypes.Structure):
    _fields_ = [("submitted", _SZ), ("completed", _SZ), ("rejected", _SZ), ("active", _SZ),
                ("queued", _SZ), ("steps", _SZ), ("bytes_in_use", _SZ), ("peak_bytes", _SZ),
                ("peak_active", _SZ), ("busy_s", ctypes.c_double)]


_lib = None


def lib():
    """Loads libbnmf_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it (python -c 'import "
                               f"__graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(LIB_PATH)
        st = ctypes.c_int
        L.bnmf_last_error.restype = ctypes.c_char_p
        L.bnmf_abi_version.restype = ctypes.c_int
        L.bnmf_init.argtypes = [ctypes.c_int]
        L.bnmf_init.restype = st
        L.bnmf_bld_mult_f64.argtypes = [_VP, _SZ, _SZ, _SZ, _D, _D, _SZ, ctypes.c_int,
                                        ctypes.c_double, _VP, ctypes.c_uint64, _VP]
        L.bnmf_bld_mult_f64.restype = st
        L.bnmf_bld_mult_f32.argtypes = L.bnmf_bld_mult_f64.argtypes
        L.bnmf_bld_mult_f32.restype = st
        # device pointers (X, alpha, beta, S0, S_out) and a hipStream_t
        L.bnmf_bld_mult_f32_noalloc.argtypes = [_VP, _SZ, _SZ, _SZ, _VP, _VP, _SZ, ctypes.c_int,
                                                ctypes.c_double, _VP, ctypes.c_uint64, _VP, _VP]
        L.bnmf_bld_mult_f32_noalloc.restype = st
        L.bnmf_bld_add_f64.argtypes = [_VP, _SZ, _SZ, _SZ, _D, _D, _SZ, ctypes.c_double, _VP,
                                       ctypes.c_uint64, _VP]
        L.bnmf_bld_add_f64.restype = st
        L.bnmf_bld_add_f32.argtypes = L.bnmf_bld_add_f64.argtypes
        L.bnmf_bld_add_f32.restype = st
        L.bnmf_log_marginal_S.argtypes = [ctypes.c_int, _VP, _SZ, _SZ, _SZ, _D, _D,
                                          ctypes.c_double, ctypes.c_double, _D]
        L.bnmf_log_marginal_S.restype = st
        L.bnmf_tensor_sums.argtypes = [ctypes.c_int, _VP, _SZ, _SZ, _SZ, _VP, _VP, _VP]
        L.bnmf_tensor_sums.restype = st
        L.bnmf_batch_create.argtypes = [ctypes.POINTER(_VP), ctypes.c_int,
                                        ctypes.POINTER(bnmf_bld_desc), _SZ, ctypes.c_int,
                                        ctypes.c_double]
        L.bnmf_batch_create.restype = st
        L.bnmf_batch_create_algo.argtypes = [ctypes.POINTER(_VP), ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(bnmf_bld_desc), _SZ, ctypes.c_int,
                                             ctypes.c_double]
        L.bnmf_batch_create_algo.restype = st
        L.bnmf_bld_appr_f64.argtypes = [_VP, _SZ, _SZ, _SZ, _D, _D, _SZ, ctypes.c_double, _VP, _VP,
                                        ctypes.c_uint64, _VP, _VP, _VP]
        L.bnmf_bld_appr_f64.restype = st
        L.bnmf_bld_appr_f32.argtypes = L.bnmf_bld_appr_f64.argtypes
        L.bnmf_bld_appr_f32.restype = st
        L.bnmf_appr_create.argtypes = [ctypes.POINTER(_VP), ctypes.c_int,
                                       ctypes.POINTER(bnmf_appr_desc), _SZ, ctypes.c_double]
        L.bnmf_appr_create.restype = st
        L.bnmf_appr_solve.argtypes = [_VP, _SZ]
        L.bnmf_appr_solve.restype = st
        L.bnmf_appr_get.argtypes = [_VP, _SZ, _VP, _VP, _VP]
        L.bnmf_appr_get.restype = st
        L.bnmf_appr_time_iteration.argtypes = [_VP, ctypes.c_int, _D]
        L.bnmf_appr_time_iteration.restype = st
        L.bnmf_appr_destroy.argtypes = [_VP]
        L.bnmf_appr_destroy.restype = None
        L.bnmf_batch_solve_tol.argtypes = [_VP, _SZ, ctypes.c_double, _SZ, ctypes.POINTER(_SZ), _D]
        L.bnmf_batch_solve_tol.restype = st
        for name in ("bnmf_batch_solve", "bnmf_batch_solve_async"):
            getattr(L, name).argtypes = [_VP, _SZ]
            getattr(L, name).restype = st
        L.bnmf_batch_synchronize.argtypes = [_VP]
        L.bnmf_batch_synchronize.restype = st
        L.bnmf_batch_get_S.argtypes = [_VP, _SZ, _VP]
        L.bnmf_batch_get_S.restype = st
        L.bnmf_batch_log_marginal.argtypes = [_VP, _D]
        L.bnmf_batch_log_marginal.restype = st
        L.bnmf_batch_time_pass.argtypes = [_VP, ctypes.c_int, _D, _D]
        L.bnmf_batch_time_pass.restype = st
        L.bnmf_batch_get_info.argtypes = [_VP, ctypes.POINTER(bnmf_batch_info)]
        L.bnmf_batch_get_info.restype = st
        L.bnmf_batch_destroy.argtypes = [_VP]
        L.bnmf_batch_destroy.restype = None
        L.bnmf_sample_X.argtypes = [ctypes.c_uint64, _SZ, _SZ, _SZ, _D, _D, ctypes.c_double,
                                    ctypes.c_double, _D]
        L.bnmf_sample_X.restype = st
        L.bnmf_sample_S.argtypes = L.bnmf_sample_X.argtypes
        L.bnmf_sample_X_gpu.argtypes = L.bnmf_sample_X.argtypes
        L.bnmf_sample_X_gpu.restype = st
        L.bnmf_sample_S_gpu.argtypes = L.bnmf_sample_X.argtypes
        L.bnmf_sample_S_gpu.restype = st
        L.bnmf_sample_S.restype = st
        L.bnmf_init_S.argtypes = [ctypes.c_int, ctypes.c_uint64, _VP, _SZ, _SZ, _SZ, _VP]
        L.bnmf_init_S.restype = st
        L.bnmf_psi_eval.argtypes = [ctypes.c_int, _D, _D, _SZ]
        L.bnmf_psi_eval.restype = st
        L.bnmf_batch_fact.argtypes = [_VP, _SZ, ctypes.c_double, _D, _D, _D]
        L.bnmf_batch_fact.restype = st
        L.bnmf_bld_fact.argtypes = [ctypes.c_int, _VP, _SZ, _SZ, _SZ, _D, _D, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, _D, _D, _D]
        L.bnmf_bld_fact.restype = st
        L.bnmf_tune_set.argtypes = [ctypes.c_int, _SZ, _SZ, ctypes.c_uint]
        L.bnmf_tune_set.restype = st
        L.bnmf_tune_get.argtypes = [ctypes.c_int, _SZ, _SZ, ctypes.POINTER(ctypes.c_uint)]
        L.bnmf_tune_get.restype = st
        L.bnmf_tune_load.argtypes = [ctypes.c_char_p]
        L.bnmf_tune_load.restype = st
        L.bnmf_tune_save.argtypes = [ctypes.c_char_p]
        L.bnmf_tune_save.restype = st
        L.bnmf_tile_rows.argtypes = [ctypes.c_int, _SZ, _SZ, _SZ]
        L.bnmf_tile_rows.restype = ctypes.c_uint
        L.bnmf_tune_source.restype = ctypes.c_char_p
        if hasattr(L, "bnmf_server_create"):
            L.bnmf_server_create.argtypes = [ctypes.POINTER(_VP), ctypes.c_int, _SZ, _SZ,
                                             ctypes.c_int, ctypes.c_double]
            L.bnmf_server_create.restype = st
            L.bnmf_server_submit.argtypes = [_VP, ctypes.POINTER(bnmf_bld_desc), _SZ,
                                             ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
            L.bnmf_server_submit.restype = st
            L.bnmf_server_poll.argtypes = [_VP, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]
            L.bnmf_server_poll.restype = st
            L.bnmf_server_wait.argtypes = [_VP, ctypes.c_uint64, ctypes.c_double]
            L.bnmf_server_wait.restype = st
            L.bnmf_server_collect.argtypes = [_VP, ctypes.c_uint64, _VP, _D,
                                              ctypes.POINTER(bnmf_request_stats)]
            L.bnmf_server_collect.restype = st
            L.bnmf_server_get_stats.argtypes = [_VP, ctypes.POINTER(bnmf_server_stats)]
            L.bnmf_server_get_stats.restype = st
            L.bnmf_server_request_bytes.argtypes = [ctypes.c_int, _SZ, _SZ, _SZ]
            L.bnmf_server_request_bytes.restype = _SZ
            L.bnmf_server_destroy.argtypes = [_VP]
            L.bnmf_server_destroy.restype = None
        _lib = L
    return _lib


class RequestRejected(RuntimeError):
    """A serving request needs more HBM than the server's whole budget."""


def _check(status: int, what: str):
    if status == OK:
        return
    msg = lib().bnmf_last_error().decode()
    if status == ERR_INVALID_ARGUMENT:
        raise ValueError(f"{what}: {msg}")
    if status == ERR_REJECTED:
        raise RequestRejected(f"{what}: {msg}")
    if status == ERR_OUT_OF_MEMORY:
        raise MemoryError(f"{what}: {msg}")
    raise RuntimeError(f"{what}: status {status}: {msg}")


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _dptr(a):
    return a.ctypes.data_as(_D)


def _np_dtype(dtype) -> np.dtype:
    dt = np.dtype(dtype)
    if dt not in (np.dtype(np.float32), np.dtype(np.float64)):
        raise ValueError("dtype must be float32 or float64")
    return dt


def _code(dt: np.dtype) -> int:
    return F32 if dt == np.float32 else F64


@dataclass
class Params:
    """alloc_model::Params: Gamma(a, b) prior on L, Dirichlet(alpha) on W's
    columns (len x), Dirichlet(beta) on H's columns (len z)."""
    a: float
    b: float
    alpha: Sequence[float]
    beta: Sequence[float]

    @classmethod
    def default(cls, shape):
        """Params(shape) (alloc_model_params.hpp:78-81): a=1, b=10, ones."""
        x, _, z = shape
        return cls(1.0, 10.0, [1.0] * x, [1.0] * z)


def init(device: int = 0):
    _check(lib().bnmf_init(int(device)), "bnmf_init")


def _check_bld_params(X: np.ndarray, z: int, params: Params):
    # check_bld_params (src/bld/util_details.hpp:20-28)
    if X.ndim != 2:
        raise ValueError("X must be a matrix")
    if len(params.alpha) != X.shape[0]:
        raise ValueError("Number of alpha parameters must be equal to number of rows of X")
    if len(params.beta) != z:
        raise ValueError("Number of beta parameters must be equal to z")


def bld_mult(X, z: int, model_params: Params, max_iter: int = 1000,
             use_psi_appr: bool = False, eps: float = 1e-50, *, S0=None, seed: int = 0,
             dtype=np.float64) -> np.ndarray:
    """bld::bld_mult: S (x, y, z) with S_ij+ ~= X after max_iter
    multiplicative updates, on the GPU. dtype float64 runs the plain FP64
    form, float32 the FP32-bulk form. S0 overrides the seeded init_S."""
    dt = _np_dtype(dtype)
    X = np.ascontiguousarray(X, dtype=dt)
    _check_bld_params(X, z, model_params)
    x, y = X.shape
    alpha = np.ascontiguousarray(model_params.alpha, dtype=np.float64)
    beta = np.ascontiguousarray(model_params.beta, dtype=np.float64)
    S0a = None
    if S0 is not None:
        S0a = np.ascontiguousarray(S0, dtype=dt)
        if S0a.shape != (x, y, z):
            raise ValueError("S0 must have shape (x, y, z)")
    S = np.empty((x, y, z), dtype=dt)
    fn = lib().bnmf_bld_mult_f32 if dt == np.float32 else lib().bnmf_bld_mult_f64
    _check(fn(_ptr(X), x, y, z, _dptr(alpha), _dptr(beta), int(max_iter), int(bool(use_psi_appr)),
              float(eps), _ptr(S0a), int(seed) & (2**64 - 1), _ptr(S)), "bld_mult")
    return S


# The reference's GPU driver has the same signature; both run on the GPU here.
bld_mult_cuda = bld_mult

ALGO_MULT = 0
ALGO_ADD = 1


def bld_add(X, z: int, model_params: Params, max_iter: int = 1000, eps: float = 1e-50, *,
            S0=None, seed: int = 0, dtype=np.float64) -> np.ndarray:
    """bld::bld_add (src/bld/bld_add.hpp:69-184): S (x, y, z) with S_ij+ = X
    after max_iter additive gradient-ascent steps on log S, on the GPU.
    Every digamma is psi_appr, as in the reference. dtype float64 or
    float32; S0 overrides the seeded initial tensor."""
    dt = _np_dtype(dtype)
    X = np.ascontiguousarray(X, dtype=dt)
    _check_bld_params(X, z, model_params)
    x, y = X.shape
    alpha = np.ascontiguousarray(model_params.alpha, dtype=np.float64)
    beta = np.ascontiguousarray(model_params.beta, dtype=np.float64)
    S0a = None
    if S0 is not None:
        S0a = np.ascontiguousarray(S0, dtype=dt)
        if S0a.shape != (x, y, z):
            raise ValueError("S0 must have shape (x, y, z)")
    S = np.empty((x, y, z), dtype=dt)
    fn = lib().bnmf_bld_add_f32 if dt == np.float32 else lib().bnmf_bld_add_f64
    _check(fn(_ptr(X), x, y, z, _dptr(alpha), _dptr(beta), int(max_iter), float(eps), _ptr(S0a),
              int(seed) & (2**64 - 1), _ptr(S)), "bld_add")
    return S


def bld_appr(X, z: int, model_params: Params, max_iter: int = 1000, eps: float = 1e-50, *,
             nu0=None, mu0=None, seed: int = 0, dtype=np.float64):
    """bld::bld_appr (src/bld/bld_appr.hpp:287-372) on the GPU: returns
    (S (x, y, z), nu (x, z), mu (z, y)). nu0 / mu0 override the seeded
    Dirichlet(1) draws of nu's rows and mu's columns."""
    dt = _np_dtype(dtype)
    X = np.ascontiguousarray(X, dtype=dt)
    _check_bld_params(X, z, model_params)
    x, y = X.shape
    alpha = np.ascontiguousarray(model_params.alpha, dtype=np.float64)
    beta = np.ascontiguousarray(model_params.beta, dtype=np.float64)
    nu0a = None if nu0 is None else np.ascontiguousarray(nu0, dtype=dt)
    mu0a = None if mu0 is None else np.ascontiguousarray(mu0, dtype=dt)
    if nu0a is not None and nu0a.shape != (x, z):
        raise ValueError("nu0 must have shape (x, z)")
    if mu0a is not None and mu0a.shape != (z, y):
        raise ValueError("mu0 must have shape (z, y)")
    S, nu, mu = np.empty((x, y, z), dtype=dt), np.empty((x, z), dtype=dt), np.empty((z, y), dtype=dt)
    fn = lib().bnmf_bld_appr_f32 if dt == np.float32 else lib().bnmf_bld_appr_f64
    _check(fn(_ptr(X), x, y, z, _dptr(alpha), _dptr(beta), int(max_iter), float(eps), _ptr(nu0a),
              _ptr(mu0a), int(seed) & (2**64 - 1), _ptr(S), _ptr(nu), _ptr(mu)), "bld_appr")
    return S, nu, mu


class ApprBatch:
    """Many bld_appr instances solved together on the GPU; the factors stay
    resident. Instances are (X, z, params, nu0, mu0, seed) tuples or objects
    with those attributes (nu0 / mu0 may be None)."""

    def __init__(self, instances, dtype=np.float64, eps: float = 1e-50):
        self.dt = _np_dtype(dtype)
        arr = (bnmf_appr_desc * max(1, len(instances)))()
        keep = []
        self.shapes = []
        for n, inst in enumerate(instances):
            X = np.ascontiguousarray(inst.X, dtype=self.dt)
            _check_bld_params(X, inst.z, inst.params)
            alpha = np.ascontiguousarray(inst.params.alpha, dtype=np.float64)
            beta = np.ascontiguousarray(inst.params.beta, dtype=np.float64)
            nu0 = getattr(inst, "nu0", None)
            mu0 = getattr(inst, "mu0", None)
            nu0 = None if nu0 is None else np.ascontiguousarray(nu0, dtype=self.dt)
            mu0 = None if mu0 is None else np.ascontiguousarray(mu0, dtype=self.dt)
            keep += [X, alpha, beta, nu0, mu0]
            d = arr[n]
            d.x, d.y, d.z = X.shape[0], X.shape[1], inst.z
            d.X = X.ctypes.data
            d.alpha, d.beta = _dptr(alpha), _dptr(beta)
            d.nu0 = None if nu0 is None else nu0.ctypes.data
            d.mu0 = None if mu0 is None else mu0.ctypes.data
            d.seed = int(getattr(inst, "seed", 0)) & (2**64 - 1)
            self.shapes.append((X.shape[0], X.shape[1], inst.z))
        h = _VP()
        _check(lib().bnmf_appr_create(ctypes.byref(h), _code(self.dt), arr, len(instances),
                                      float(eps)), "appr_create")
        del keep
        self._h = h

    def solve(self, max_iter: int):
        _check(lib().bnmf_appr_solve(self._h, int(max_iter)), "appr_solve")

    def get(self, idx: int):
        x, y, z = self.shapes[idx]
        S = np.empty((x, y, z), dtype=self.dt)
        nu, mu = np.empty((x, z), dtype=self.dt), np.empty((z, y), dtype=self.dt)
        _check(lib().bnmf_appr_get(self._h, int(idx), _ptr(S), _ptr(nu), _ptr(mu)), "appr_get")
        return S, nu, mu

    def time_iteration(self, reps: int = 10) -> float:
        out = np.zeros(1)
        _check(lib().bnmf_appr_time_iteration(self._h, int(reps), _dptr(out)), "appr_time")
        return float(out[0])

    def close(self):
        if getattr(self, "_h", None):
            lib().bnmf_appr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def log_marginal_S(S, model_params: Params) -> float:
    """alloc_model::log_marginal_S in FP64 on the GPU."""
    S = np.asarray(S)
    dt = np.dtype(np.float32) if S.dtype == np.float32 else np.dtype(np.float64)
    S = np.ascontiguousarray(S, dtype=dt)
    x, y, z = S.shape
    if len(model_params.alpha) != x:
        raise ValueError("Number of alpha parameters must be equal to S.dimension(0)")
    if len(model_params.beta) != z:
        raise ValueError("Number of beta parameters must be equal to S.dimension(2)")
    alpha = np.ascontiguousarray(model_params.alpha, dtype=np.float64)
    beta = np.ascontiguousarray(model_params.beta, dtype=np.float64)
    out = np.zeros(1)
    _check(lib().bnmf_log_marginal_S(_code(dt), _ptr(S), x, y, z, _dptr(alpha), _dptr(beta),
                                     float(model_params.a), float(model_params.b), _dptr(out)),
           "log_marginal_S")
    return float(out[0])


def tensor_sums(S):
    """cuda::tensor_sums: (S_pjk (y,z), S_ipk (x,z), S_ijp (x,y))."""
    S = np.asarray(S)
    dt = np.dtype(np.float32) if S.dtype == np.float32 else np.dtype(np.float64)
    S = np.ascontiguousarray(S, dtype=dt)
    x, y, z = S.shape
    pjk, ipk, ijp = (np.empty((y, z), dt), np.empty((x, z), dt), np.empty((x, y), dt))
    _check(lib().bnmf_tensor_sums(_code(dt), _ptr(S), x, y, z, _ptr(pjk), _ptr(ipk), _ptr(ijp)),
           "tensor_sums")
    return pjk, ipk, ijp


def sample_X(seed: int, shape, model_params: Params) -> np.ndarray:
    """X = S_ij+ of an allocation-model sample (bnmf_priors + sample_S)."""
    x, y, z = shape
    alpha = np.ascontiguousarray(model_params.alpha, dtype=np.float64)
    beta = np.ascontiguousarray(model_params.beta, dtype=np.float64)
    X = np.empty((x, y))
    _check(lib().bnmf_sample_X(int(seed), x, y, z, _dptr(alpha), _dptr(beta),
                               float(model_params.a), float(model_params.b), _dptr(X)), "sample_X")
    return X


def sample_X_gpu(seed: int, shape, model_params: Params) -> np.ndarray:
    """sample_X with the Poisson draws on the GPU (same streams)."""
    x, y, z = shape
    alpha = np.ascontiguousarray(model_params.alpha, dtype=np.float64)
    beta = np.ascontiguousarray(model_params.beta, dtype=np.float64)
    X = np.empty((x, y))
    _check(lib().bnmf_sample_X_gpu(int(seed), x, y, z, _dptr(alpha), _dptr(beta),
                                   float(model_params.a), float(model_params.b), _dptr(X)),
           "sample_X_gpu")
    return X


def sample_S_gpu(seed: int, shape, model_params: Params) -> np.ndarray:
    """sample_S with the Poisson draws on the GPU (same streams)."""
    x, y, z = shape
    alpha = np.ascontiguousarray(model_params.alpha, dtype=np.float64)
    beta = np.ascontiguousarray(model_params.beta, dtype=np.float64)
    S = np.empty((x, y, z))
    _check(lib().bnmf_sample_S_gpu(int(seed), x, y, z, _dptr(alpha), _dptr(beta),
                                   float(model_params.a), float(model_params.b), _dptr(S)),
           "sample_S_gpu")
    return S


def sample_S(seed: int, shape, model_params: Params) -> np.ndarray:
    x, y, z = shape
    alpha = np.ascontiguousarray(model_params.alpha, dtype=np.float64)
    beta = np.ascontiguousarray(model_params.beta, dtype=np.float64)
    S = np.empty((x, y, z))
    _check(lib().bnmf_sample_S(int(seed), x, y, z, _dptr(alpha), _dptr(beta),
                               float(model_params.a), float(model_params.b), _dptr(S)), "sample_S")
    return S


def init_S(seed: int, X, z: int, dtype=np.float64) -> np.ndarray:
    dt = _np_dtype(dtype)
    X = np.ascontiguousarray(X, dtype=dt)
    x, y = X.shape
    S = np.empty((x, y, z), dtype=dt)
    _check(lib().bnmf_init_S(_code(dt), int(seed), _ptr(X), x, y, z, _ptr(S)), "init_S")
    return S


def bld_fact(S, model_params: Params, eps: float = 1e-50):
    """bld::bld_fact (src/bld/bld_fact.hpp:36-80): W (x, z), H (z, y), L (y,)
    from S, with the axis sums and normalisation on the GPU."""
    S = np.asarray(S)
    dt = np.dtype(np.float32) if S.dtype == np.float32 else np.dtype(np.float64)
    S = np.ascontiguousarray(S, dtype=dt)
    x, y, z = S.shape
    alpha = np.ascontiguousarray(model_params.alpha, dtype=np.float64)
    beta = np.ascontiguousarray(model_params.beta, dtype=np.float64)
    if alpha.shape != (x,) or beta.shape != (z,):
        raise ValueError("alpha must have x entries and beta z entries")
    W, H, L = np.empty((x, z)), np.empty((z, y)), np.empty(y)
    _check(lib().bnmf_bld_fact(_code(dt), _ptr(S), x, y, z, _dptr(alpha), _dptr(beta),
                               float(model_params.a), float(model_params.b), float(eps),
                               _dptr(W), _dptr(H), _dptr(L)), "bld_fact")
    return W, H, L


def psi_eval(kind: int, x) -> np.ndarray:
    """Device digamma forms (0: FP32 bulk, 1: exact FP64, 2: psi_appr FP64,
    3: FP32 psi(s+1) of the update passes evaluated at x = s+1)."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    out = np.empty_like(x)
    _check(lib().bnmf_psi_eval(int(kind), _dptr(x), _dptr(out), x.size), "psi_eval")
    return out


@dataclass
class Instance:
    """One BLD problem of a batch: X (x, y), depth z, model params, and an
    optional S0 (else drawn as init_S with `seed`)."""
    X: np.ndarray
    z: int
    params: Params
    S0: Optional[np.ndarray] = None
    seed: int = 0
    _keep: list = field(default_factory=list, repr=False)


def _descs(instances: Sequence[Instance], dt: np.dtype):
    arr = (bnmf_bld_desc * max(1, len(instances)))()
    keep = []
    for n, inst in enumerate(instances):
        X = np.ascontiguousarray(inst.X, dtype=dt)
        _check_bld_params(X, inst.z, inst.params)
        alpha = np.ascontiguousarray(inst.params.alpha, dtype=np.float64)
        beta = np.ascontiguousarray(inst.params.beta, dtype=np.float64)
        S0 = None if inst.S0 is None else np.ascontiguousarray(inst.S0, dtype=dt)
        if S0 is not None and S0.shape != (X.shape[0], X.shape[1], inst.z):
            raise ValueError("S0 must have shape (x, y, z)")
        keep += [X, alpha, beta, S0]
        d = arr[n]
        d.x, d.y, d.z = X.shape[0], X.shape[1], inst.z
        d.X = X.ctypes.data
        d.alpha = _dptr(alpha)
        d.beta = _dptr(beta)
        d.a, d.b = float(inst.params.a), float(inst.params.b)
        d.S0 = None if S0 is None else S0.ctypes.data
        d.seed = int(inst.seed) & (2**64 - 1)
    return arr, keep


class Batch:
    """Many independent instances solved together on the GPU (batched fused
    pass over a device work list, per-iteration chain replayed from a
    hipGraph). Inputs are uploaded once; S stays resident in HBM."""

    def __init__(self, instances: Sequence[Instance], dtype=np.float32,
                 use_psi_appr: bool = True, eps: float = 1e-50, algo: int = ALGO_MULT):
        self.dt = _np_dtype(dtype)
        self.shapes = [(np.shape(i.X)[0], np.shape(i.X)[1], i.z) for i in instances]
        arr, keep = _descs(instances, self.dt)
        h = _VP()
        _check(lib().bnmf_batch_create_algo(ctypes.byref(h), _code(self.dt), int(algo), arr,
                                            len(instances), int(bool(use_psi_appr)), float(eps)),
               "batch_create")
        del keep
        self._h = h

    def solve(self, max_iter: int):
        _check(lib().bnmf_batch_solve(self._h, int(max_iter)), "batch_solve")

    def solve_async(self, max_iter: int):
        _check(lib().bnmf_batch_solve_async(self._h, int(max_iter)), "batch_solve_async")

    def solve_tol(self, max_iter: int, tol: float, check_every: int = 10):
        """Solve from S0 with the FP64 log-marginal convergence check: stops
        once every instance's relative log-marginal change over check_every
        iterations is <= tol. Returns (iterations run, log-marginals)."""
        n = ctypes.c_size_t(0)
        lm = np.zeros(max(1, len(self.shapes)))
        _check(lib().bnmf_batch_solve_tol(self._h, int(max_iter), float(tol), int(check_every),
                                          ctypes.byref(n), _dptr(lm)), "batch_solve_tol")
        return int(n.value), lm[:len(self.shapes)]

    def synchronize(self):
        _check(lib().bnmf_batch_synchronize(self._h), "batch_synchronize")

    def S(self, idx: int) -> np.ndarray:
        out = np.empty(self.shapes[idx], dtype=self.dt)
        _check(lib().bnmf_batch_get_S(self._h, int(idx), _ptr(out)), "batch_get_S")
        return out

    def log_marginal(self) -> np.ndarray:
        out = np.zeros(max(1, len(self.shapes)))
        _check(lib().bnmf_batch_log_marginal(self._h, _dptr(out)), "batch_log_marginal")
        return out[:len(self.shapes)]

    def fact(self, idx: int, eps: float = 1e-50):
        """bld::bld_fact of instance idx's current S: (W (x, z), H (z, y), L (y,))."""
        x, y, z = self.shapes[idx]
        W, H, L = np.empty((x, z)), np.empty((z, y)), np.empty(y)
        _check(lib().bnmf_batch_fact(self._h, int(idx), float(eps), _dptr(W), _dptr(H), _dptr(L)),
               "batch_fact")
        return W, H, L

    def time_pass(self, reps: int = 20):
        p, it = np.zeros(1), np.zeros(1)
        _check(lib().bnmf_batch_time_pass(self._h, int(reps), _dptr(p), _dptr(it)), "time_pass")
        return float(p[0]), float(it[0])

    def info(self) -> dict:
        inf = bnmf_batch_info()
        _check(lib().bnmf_batch_get_info(self._h, ctypes.byref(inf)), "batch_get_info")
        return {k: getattr(inf, k) for k, _ in bnmf_batch_info._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().bnmf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Server:
    """Continuous-batching solve server on the GPU (C ABI bnmf_server_*).

    Requests are admitted in arrival order while their HBM footprint fits in
    `hbm_budget` bytes; active requests advance together and new ones join at
    the next scheduling step. S stays in HBM until collected."""

    def __init__(self, hbm_budget: int, dtype=np.float32, use_psi_appr: bool = True,
                 eps: float = 1e-50, max_active: int = 0):
        self.dt = _np_dtype(dtype)
        h = _VP()
        _check(lib().bnmf_server_create(ctypes.byref(h), _code(self.dt), int(hbm_budget),
                                        int(max_active), int(bool(use_psi_appr)), float(eps)),
               "server_create")
        self._h = h
        self._shapes = {}

    @staticmethod
    def request_bytes(shape, dtype=np.float32) -> int:
        x, y, z = shape
        return int(lib().bnmf_server_request_bytes(_code(_np_dtype(dtype)), x, y, z))

    def submit(self, inst: Instance, max_iter: int, log_marginal: bool = False) -> int:
        arr, keep = _descs([inst], self.dt)
        t = ctypes.c_uint64()
        _check(lib().bnmf_server_submit(self._h, arr, int(max_iter), int(bool(log_marginal)),
                                        ctypes.byref(t)), "server_submit")
        del keep
        x, y = np.shape(inst.X)
        self._shapes[t.value] = (x, y, inst.z)
        return t.value

    def poll(self, ticket: int) -> bool:
        d = ctypes.c_int()
        _check(lib().bnmf_server_poll(self._h, int(ticket), ctypes.byref(d)), "server_poll")
        return bool(d.value)

    def wait(self, ticket: int, timeout: float = -1.0) -> int:
        return int(lib().bnmf_server_wait(self._h, int(ticket), float(timeout)))

    def collect(self, ticket: int, want_S: bool = True):
        """Returns (S or None, log_marginal, stats dict); raises on failure."""
        shape = self._shapes.pop(int(ticket), None)
        S = np.empty(shape, dtype=self.dt) if (want_S and shape is not None) else None
        lm = np.zeros(1)
        st = bnmf_request_stats()
        _check(lib().bnmf_server_collect(self._h, int(ticket), _ptr(S), _dptr(lm),
                                         ctypes.byref(st)), "server_collect")
        stats = {k: getattr(st, k) for k, _ in bnmf_request_stats._fields_}
        return S, float(lm[0]), stats

    def stats(self) -> dict:
        st = bnmf_server_stats()
        _check(lib().bnmf_server_get_stats(self._h, ctypes.byref(st)), "server_get_stats")
        return {k: getattr(st, k) for k, _ in bnmf_server_stats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().bnmf_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# --------------------------------------------------------------- autotuner
# Shape classes of the tile table (include/bnmf_hip.h): z classes end at
# these depths; size classes split fibers = x*y at 2^14, 2^16, 2^18, 2^20.
TUNE_Z_CLASSES = (4, 8, 12, 16, 24, 32, 48, 64)
TUNE_FIBER_CLASSES = ((1 << 10, 1 << 14), (1 << 14, 1 << 16), (1 << 16, 1 << 18),
                      (1 << 18, 1 << 20), (1 << 20, 1 << 22))


def tune_set(dtype, z: int, fibers: int, rows: int):
    """Sets the tile height of the shape class of (z, fibers); 0 = rule."""
    _check(lib().bnmf_tune_set(_code(_np_dtype(dtype)), int(z), int(fibers), int(rows)), "tune_set")


def tune_get(dtype, z: int, fibers: int) -> int:
    r = ctypes.c_uint(0)
    _check(lib().bnmf_tune_get(_code(_np_dtype(dtype)), int(z), int(fibers), ctypes.byref(r)),
           "tune_get")
    return int(r.value)


def tune_load(path: str):
    _check(lib().bnmf_tune_load(os.fsencode(path)), "tune_load")


def tune_save(path: str):
    _check(lib().bnmf_tune_save(os.fsencode(path)), "tune_save")


def tune_source() -> str:
    return lib().bnmf_tune_source().decode()


def tile_rows(dtype, x: int, y: int, z: int) -> int:
    """Rows per pass tile an instance of this shape gets (table or rule)."""
    return int(lib().bnmf_tile_rows(_code(_np_dtype(dtype)), int(x), int(y), int(z)))


def _tune_class(z: int, fibers: int):
    zc = next(c for c, hi in enumerate(TUNE_Z_CLASSES) if z <= hi or c == len(TUNE_Z_CLASSES) - 1)
    fc = next(c for c, (lo, hi) in enumerate(TUNE_FIBER_CLASSES)
              if fibers < hi or c == len(TUNE_FIBER_CLASSES) - 1)
    return zc, fc


def autotune_workload(dtype, instances: Sequence[Instance], candidates=(16, 32, 64),
                      reps: int = 10, trials: int = 2, refine: bool = True,
                      use_psi_appr: bool = True, log=None) -> dict:
    """Chooses tile heights for a workload solved as one batch, timing whole
    iterations of the whole batch (graph-launched, as solves run them), so
    the effects of running all shape classes together are measured: first
    one height for every class (best of `candidates`), then, if `refine`,
    each class (largest first) tries the other candidates with the rest
    fixed and keeps a change that makes the iteration at least 1 % faster.
    Fills the tile table; returns {"rows": {(zc, fc): rows}, "iter_ms": t,
    "evaluations": n}."""
    dt = _np_dtype(dtype)
    groups = {}
    for inst in instances:
        x, y = np.shape(inst.X)
        if x * y * inst.z:
            groups.setdefault(_tune_class(inst.z, x * y), []).append(inst)
    reps_of = {k: (v[0].z, np.shape(v[0].X)[0] * np.shape(v[0].X)[1]) for k, v in groups.items()}
    evals = 0

    def measure(choice):
        nonlocal evals
        for k, rows in choice.items():
            tune_set(dt, reps_of[k][0], reps_of[k][1], rows)
        with Batch(instances, dtype=dt, use_psi_appr=use_psi_appr) as b:
            b.solve(2)
            t = min(b.time_pass(reps)[1] for _ in range(trials))
        evals += 1
        return t

    best_t, best = None, None
    for c in candidates:
        choice = {k: c for k in groups}
        t = measure(choice)
        if log:
            log(f"[autotune] all classes {c} rows: {t * 1e3:.1f} us per iteration")
        if best_t is None or t < best_t:
            best_t, best = t, choice
    if refine:
        order = sorted(groups, key=lambda k: -sum(np.size(i.X) * i.z for i in groups[k]))
        for k in order:
            for c in candidates:
                if c == best[k]:
                    continue
                trial = dict(best)
                trial[k] = c
                t = measure(trial)
                if t < best_t * 0.99:
                    if log:
                        log(f"[autotune] class z{k[0]} f{k[1]}: {best[k]} -> {c} rows, "
                            f"{best_t * 1e3:.1f} -> {t * 1e3:.1f} us")
                    best_t, best = t, trial
    for k, rows in best.items():
        tune_set(dt, reps_of[k][0], reps_of[k][1], rows)
    return {"rows": best, "iter_ms": best_t, "evaluations": evals}


def autotune(dtype=np.float32, instances: Optional[Sequence[Instance]] = None,
             z_classes=TUNE_Z_CLASSES, fiber_classes=TUNE_FIBER_CLASSES,
             candidates=(16, 32, 64, 128, 256), reps: int = 10, trials: int = 3,
             target_elements: int = 1 << 26, max_instances: int = 256,
             path: Optional[str] = None, seed: int = 0x7A7E, use_psi_appr: bool = True,
             log=None) -> dict:
    """Times the fused pass for each tile height candidate, per shape class,
    and keeps the fastest in the tile table (saved to `path` if given).

    With `instances` (a workload), each class present in it is timed on its
    own instances, solved together as a batch. Without, each class of
    `z_classes` x `fiber_classes` gets a synthetic workload: n instances of
    x = y = sqrt(geometric mid of the fiber range) and z = the class's
    largest depth, n chosen so the batch holds about `target_elements`
    elements, X from the allocation model (a = 40, b = 1).
    A candidate's time is the best of `trials` timings of `reps` passes
    (graph-launched, as solves run them); candidates that tile every
    instance of the class as an earlier one did are skipped.
    Returns {(zc, fc): {"rows": best, "pass_ms": {rows: ms}, ...}}."""
    dt = _np_dtype(dtype)
    groups = {}
    if instances is not None:
        for inst in instances:
            x, y = np.shape(inst.X)
            if x * y * inst.z:
                groups.setdefault(_tune_class(inst.z, x * y), []).append(inst)
    else:
        rng = np.random.default_rng(seed)
        for z in z_classes:
            for lo, hi in fiber_classes:
                side = max(1, int(round(np.sqrt(np.sqrt(lo * (hi - 1))))))
                n = int(min(max_instances, max(1, target_elements // (side * side * z))))
                p = Params(40.0, 1.0, [1.0] * side, [1.0] * z)
                groups[_tune_class(z, side * side)] = [
                    Instance(X=sample_X(int(rng.integers(1 << 62)), (side, side, z), p).astype(dt),
                             z=z, params=p, seed=int(rng.integers(1 << 62))) for _ in range(n)]
    results = {}
    for key in sorted(groups):
        insts = groups[key]
        shapes = [(np.shape(i.X)[0], np.shape(i.X)[1], i.z) for i in insts]
        z0, f0 = shapes[0][2], shapes[0][0] * shapes[0][1]
        timed, seen = {}, set()
        for cand in candidates:
            tune_set(dt, z0, f0, cand)
            tiling = tuple(tile_rows(dt, *sh) for sh in shapes)
            if tiling in seen:
                continue
            seen.add(tiling)
            with Batch(insts, dtype=dt, use_psi_appr=use_psi_appr) as b:
                b.solve(2)
                timed[cand] = min(b.time_pass(reps)[0] for _ in range(trials))
        best = min(timed, key=timed.get)
        tune_set(dt, z0, f0, best)
        results[key] = {"rows": best, "pass_ms": timed, "shape": shapes[0], "instances": len(insts)}
        if log:
            log(f"[autotune] class z{key[0]} f{key[1]} ({len(insts)} instances, e.g. "
                f"{shapes[0]}): " + " ".join(f"{c}:{t * 1e3:.0f}us" for c, t in timed.items()) +
                f" -> {best}")
    if path:
        tune_save(path)
    return results
