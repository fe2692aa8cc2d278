"""The no-allocation entry point bnmf_bld_mult_f32_noalloc, called through
the C ABI on device buffers the test allocates with the HIP runtime the
library itself uses: its S must equal bnmf_bld_mult_f32's bit for bit (same
kernels, same tiles), on consecutive calls that reuse the library's scratch
with a smaller instance second, with a seeded start, on the generic pass
(beta < 1) and on the wide pass (z > 64). An instance whose workspace
exceeds the scratch is rejected."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip_runtime(hip):
    """The libamdhip64 the library is linked against (the copy already mapped
    into this process), so buffers live in the library's own runtime."""
    hip.lib()
    with open("/proc/self/maps") as fh:
        paths = {line.split()[-1] for line in fh if "libamdhip64.so" in line}
    rt = ctypes.CDLL(sorted(paths)[0])
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    return rt


class _Dev:
    """A device copy of a numpy array (hipMalloc'd by the test, the caller)."""

    def __init__(self, rt, a=None, nbytes=None):
        self.rt = rt
        self.nbytes = a.nbytes if a is not None else nbytes
        self.ptr = ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 16)) == 0
        if a is not None:
            assert rt.hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, 1) == 0  # host to device

    def get(self, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert self.rt.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        self.rt.hipFree(self.ptr)


def _problem(hip, x, y, z, seed, beta=1.0):
    p = hip.Params(40.0, 1.0, [1.0] * x, [beta] * z)
    X = hip.sample_X(seed, (x, y, z), hip.Params(40.0, 1.0, [1.0] * x, [1.0] * z)).astype(np.float32)
    S0 = hip.init_S(seed + 1, X, z, dtype=np.float32)
    return p, X, S0


def _noalloc(hip, X, z, p, iters, S0=None, seed=0):
    rt = _hip_runtime(hip)
    x, y = X.shape
    dX = _Dev(rt, np.ascontiguousarray(X, dtype=np.float32))
    da = _Dev(rt, np.ascontiguousarray(p.alpha, dtype=np.float64))
    db = _Dev(rt, np.ascontiguousarray(p.beta, dtype=np.float64))
    dS0 = None if S0 is None else _Dev(rt, np.ascontiguousarray(S0, dtype=np.float32))
    dS = _Dev(rt, nbytes=x * y * z * 4)
    print(f"[no_alloc] {x}x{y}x{z}, {iters} iterations, S0 {'given' if S0 is not None else 'drawn'}",
          flush=True)
    st = hip.lib().bnmf_bld_mult_f32_noalloc(dX.ptr, x, y, z, da.ptr, db.ptr, iters, 1, 1e-50,
                                             None if dS0 is None else dS0.ptr, seed, dS.ptr, None)
    assert st == hip.OK, hip.lib().bnmf_last_error()
    out = dS.get((x, y, z), np.float32)  # hipMemcpy waits for the default stream
    for d in (dX, da, db, dS0, dS):
        if d is not None:
            d.free()
    return out


def test_noalloc_equals_allocating_entry_point_twice(hip):
    p, X, S0 = _problem(hip, 300, 200, 20, seed=5)
    got = _noalloc(hip, X, 20, p, 30, S0=S0)
    want = hip.bld_mult(X, 20, p, 30, True, S0=S0, dtype=np.float32)
    assert np.array_equal(got, want)
    # a second, smaller call: the scratch keeps the first call's contents
    p2, X2, S02 = _problem(hip, 50, 60, 7, seed=6)
    got2 = _noalloc(hip, X2, 7, p2, 25, S0=S02)
    want2 = hip.bld_mult(X2, 7, p2, 25, True, S0=S02, dtype=np.float32)
    assert np.array_equal(got2, want2)


def test_noalloc_seeded_generic_and_deep(hip):
    # S0 drawn on the device from the seed
    p, X, _ = _problem(hip, 120, 90, 13, seed=7)
    got = _noalloc(hip, X, 13, p, 20, seed=4321)
    want = hip.bld_mult(X, 13, p, 20, True, seed=4321, dtype=np.float32)
    assert np.array_equal(got, want)
    # beta < 1: the generic pass (the device-side check picks it)
    pg, Xg, S0g = _problem(hip, 70, 130, 12, seed=8, beta=0.7)
    got = _noalloc(hip, Xg, 12, pg, 15, S0=S0g)
    want = hip.bld_mult(Xg, 12, pg, 15, True, S0=S0g, dtype=np.float32)
    assert np.array_equal(got, want)
    # z > 64: the wide pass
    pd, Xd, S0d = _problem(hip, 40, 50, 80, seed=9)
    got = _noalloc(hip, Xd, 80, pd, 12, S0=S0d)
    want = hip.bld_mult(Xd, 80, pd, 12, True, S0=S0d, dtype=np.float32)
    assert np.array_equal(got, want)
    # zero iterations return the start
    got = _noalloc(hip, Xd, 80, pd, 0, S0=S0d)
    assert np.array_equal(got, S0d)


def test_noalloc_rejects_instances_beyond_the_scratch(hip):
    lib = hip.lib()
    fake = 1 << 20  # aligned, never dereferenced: the size check comes first
    st = lib.bnmf_bld_mult_f32_noalloc(fake, 4096, 4096, 64, fake, fake, 5, 1, 1e-50, fake, 0, fake, None)
    assert st == hip.ERR_INVALID_ARGUMENT
    assert b"32 MiB" in lib.bnmf_last_error()
