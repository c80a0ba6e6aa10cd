"""Device covariance generators (dbg_cov_build) against the definitions of
the reference's simulation scripts (scripts/simulations/simulations_functions.jl),
and a plan built from a covariance that never leaves the device."""
import ctypes

import numpy as np
import pytest
from scipy.special import gamma, kv

from dbgenzpaper_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_fixed_dense_and_sparse(dbg):
    n = 50
    S = dbg.covariance("fixed_dense", n, jitter=0.7)
    assert np.array_equal(S, np.ones((n, n)) + 0.7 * np.eye(n))
    for n_sp in (0, 1, 4):
        M = np.diag(np.full(n, n_sp / 2))   # :33-41, as written
        for i in range(1, n_sp + 1):
            M += np.diag(np.ones(n - i), i)
        M = M + M.T
        assert np.array_equal(dbg.covariance("fixed_sparse", n, n_sp=n_sp), M)


@pytest.mark.parametrize("nu,dist,sigma2", [(1.5, 3.0, 1.0), (0.5, 12.0, 2.0), (2.5, 0.4, 0.5)])
def test_matern(dbg, nu, dist, sigma2):
    """mattern_cov (:56-79) through SciPy's Bessel K: relative 1e-13."""
    n = 120
    S = dbg.covariance("matern", n, nu=nu, dist=dist, sigma2=sigma2)
    d = np.abs(np.subtract.outer(np.arange(n), np.arange(n))).astype(float)
    z = np.sqrt(2 * nu) * d / dist
    with np.errstate(invalid="ignore", divide="ignore"):
        ref = sigma2 * 2 ** (1 - nu) / gamma(nu) * z ** nu * kv(nu, z)
    ref[d == 0] = sigma2
    assert np.allclose(S, ref, rtol=1e-13, atol=1e-300)
    assert np.array_equal(S, S.T)


def test_lowrank_factor_model(dbg):
    """B B' + diag(psi) and its standardization, as the factor-model workload
    builds them on the host: relative 1e-13 (dsyrk sums in another order)."""
    n, k = 300, 8
    rng = np.random.default_rng(3)
    B = rng.standard_normal((n, k))
    psi = rng.uniform(0.2, 1.0, n)
    S = dbg.covariance("lowrank", n, B=B, psi=psi)
    ref = B @ B.T + np.diag(psi)
    assert np.allclose(S, ref, rtol=1e-13, atol=1e-13) and np.array_equal(S, S.T)
    Ss = dbg.covariance("lowrank", n, B=B, psi=psi, standardize=True)
    assert np.allclose(Ss, W._standardize(ref), rtol=1e-13, atol=1e-14)
    assert np.all(np.diag(Ss) == 1.0)


def _hip():
    """The HIP runtime libdbgenz.so is linked against (the system ROCm one; the
    test allocates device memory through it rather than through torch, which
    ships a runtime of its own)."""
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def test_plan_from_device_covariance(dbg):
    """A factor-model covariance built into device memory goes straight to
    dbg_plan_create; the estimate equals the one from the same matrix copied
    to the host, bit for bit."""
    from dbgenzpaper_amd import _lib

    L = _lib.lib()
    hip = _hip()
    n, k, m, R = 96, 8, 2048, 3
    rng = np.random.default_rng(5)
    B = np.asfortranarray(rng.standard_normal((n, k)))
    psi = rng.uniform(0.2, 1.0, n)
    dptr = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dptr), n * n * 8) == 0
    try:
        params = np.zeros(3)
        _lib.check(L.dbg_cov_build(_lib.COV_LOWRANK, n, _lib.ptr(params, ctypes.c_double), k, B.ctypes.data,
                                   psi.ctypes.data, 1, dptr, n, 1))
        Sd = np.empty((n, n), order="F")
        assert hip.hipMemcpy(Sd.ctypes.data, dptr, n * n * 8, 2) == 0  # hipMemcpyDeviceToHost
        Sh = dbg.covariance("lowrank", n, B=B, psi=psi, standardize=True)
        assert np.array_equal(Sd, Sh)
        a = np.full(n, -np.inf)
        b = np.full(n, 1.0)
        shifts = np.asfortranarray(rng.random((n - 1, R)))

        def estimate(sigma_ptr):
            o = _lib.Opts()
            L.dbg_opts_default(ctypes.byref(o), _lib.PREC_FP64)
            o.m, o.n_reps, o.max_pts, o.max_abs_err = m, R, m, 0.0
            h = ctypes.c_void_p()
            _lib.check(L.dbg_plan_create(ctypes.byref(h), ctypes.byref(o), n, sigma_ptr, n,
                                         _lib.ptr(a, ctypes.c_double), _lib.ptr(b, ctypes.c_double)))
            try:
                _lib.check(L.dbg_plan_set_shifts(h, shifts.ctypes.data))
                r = _lib.Result()
                _lib.check(L.dbg_qmc_pnorm(h, ctypes.byref(r)))
                return r.value, r.err3
            finally:
                L.dbg_plan_destroy(h)

        vd = estimate(ctypes.cast(dptr, ctypes.POINTER(ctypes.c_double)))
        vh = estimate(_lib.ptr(Sh, ctypes.c_double))
        print("device-built covariance: P =", vd, flush=True)
        assert vd == vh and 0 < vd[0] < 1
    finally:
        hip.hipFree(dptr)


def test_leading_dimension_padding(dbg):
    """dbg_cov_build into host memory with lds > n: the n x n block equals the
    unpadded matrix, the padding rows are untouched."""
    from dbgenzpaper_amd import _lib

    L = _lib.lib()
    n, pad = 64, 5
    S = np.full((n + pad, n), 9.0, order="F")
    params = np.array([1.5, 4.0, 2.0])
    _lib.check(L.dbg_cov_build(_lib.COV_MATERN, n, _lib.ptr(params, ctypes.c_double), 0, None, None, 0,
                               S.ctypes.data, n + pad, 0))
    assert np.array_equal(S[:n, :], dbg.covariance("matern", n, nu=1.5, dist=4.0, sigma2=2.0))
    assert np.all(S[n:, :] == 9.0)
