"""The allocation-free factorization entry point (no reference counterpart)
against dbg_chol_genz_f64 on the same inputs: bit for bit, on two calls in a
row, the second with a smaller matrix (the fixed scratch keeps the first
call's contents and must be cleared by the call itself)."""
import ctypes

import numpy as np
import pytest

from tests._helpers import random_spd

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        self.h = h
        self.bufs = []

    def put(self, arr):
        arr = np.ascontiguousarray(arr) if arr.ndim == 1 else np.asfortranarray(arr)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(arr.nbytes, 16)) == 0
        assert self.h.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0  # host to device
        self.bufs.append(p)
        return p

    def get(self, p, like):
        out = np.empty_like(like, order="F")
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _case(n, seed):
    C = random_spd(n, seed)
    rng = np.random.default_rng(seed + 1)
    return C, -rng.uniform(0.5, 2.0, n), rng.uniform(0.5, 2.0, n)


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_genz_matches_dbg_chol_genz(dbg, order):
    from dbgenzpaper_amd import _lib

    L = _lib.lib()
    hip = _Hip()
    try:
        for n, seed in ((300, 91), (77, 92)):
            C, a, b = _case(n, seed)
            print(f"noalloc genz n={n} seed={seed}", flush=True)
            F = dbg.cholesky_genz(C.copy(order="F"), a.copy(), b.copy(), 2 ** 6, 2 ** 9)
            dM, da, db = hip.put(C), hip.put(a), hip.put(b)
            dperm = hip.put(np.zeros(n, np.int32))
            dscale = hip.put(np.zeros(n))
            drank = hip.put(np.zeros(1, np.int32))
            _lib.check(L.dbg_chol_genz_f64_dev_noalloc(dM, n, n, da, db, 2 ** 6, 2 ** 9, dperm, dscale, drank, None))
            assert hip.h.hipDeviceSynchronize() == 0
            U = hip.get(dM, C)
            assert int(hip.get(drank, np.zeros(1, np.int32))[0]) == F.rank == n
            assert np.array_equal(hip.get(dperm, np.zeros(n, np.int32)), F.perm)
            assert np.array_equal(np.triu(U), np.triu(F.U)) and np.array_equal(np.tril(U, -1), np.zeros((n, n)))
            assert np.array_equal(hip.get(da, a), F.a) and np.array_equal(hip.get(db, b), F.b)
            assert np.array_equal(hip.get(dscale, a), F.scale_vec)
            hip.free()
    finally:
        hip.free()


def test_noalloc_limits(dbg):
    from dbgenzpaper_amd import _lib

    L = _lib.lib()
    p = ctypes.c_void_p(8)  # never dereferenced: the arguments are refused first
    assert L.dbg_chol_genz_f64_dev_noalloc(p, 600, 600, p, p, 64, 1024, p, p, p, None) == _lib.DBG_ERR_ARG
    assert L.dbg_chol_genz_f64_dev_noalloc(p, 300, 300, p, p, 64, 256, p, p, p, None) == _lib.DBG_ERR_ARG
    assert L.dbg_chol_genz_f64_dev_noalloc(p, 10, 20, p, p, 64, 512, p, p, p, None) == _lib.DBG_ERR_ARG
