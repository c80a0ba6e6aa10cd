"""The mixed walk's FP32 tail functions (tailq.h) evaluated on the GPU, where
they use the device's approximate v_rcp_f32, v_rsq_f32, v_exp_f32 and v_log_f32,
against 40-digit values: the tail mass erfc(z)/2 = m 2^-k over z in [0, 30]
and the quantile Phi^{-1}(t 2^-k) from p = 1/2 down to 2^-3000, with both
sides of the quantile's region switch.  Same bounds as the host-build
tests: 5e-7 relative for the mass, 4 FP32 ulps of max(|y|, 1) for the
quantile."""
import ctypes
import os

import numpy as np
import pytest

from tests._helpers import normal_quantile_mp

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "devtail", "libdevtail.so")


def _lib():
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB)
    fp, ip = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)
    L.devtail_eval.argtypes = [fp, fp, ip, ctypes.c_int, fp, ip, fp]
    L.devtail_eval.restype = ctypes.c_int
    return L


def _eval(z, t, tk):
    n = len(z)
    assert len(t) == n and len(tk) == n
    z = np.ascontiguousarray(z, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    tk = np.ascontiguousarray(tk, np.int32)
    m = np.empty(n, np.float32)
    k = np.empty(n, np.int32)
    y = np.empty(n, np.float32)
    fp, ip = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)
    rc = _lib().devtail_eval(z.ctypes.data_as(fp), t.ctypes.data_as(fp), tk.ctypes.data_as(ip), n,
                             m.ctypes.data_as(fp), k.ctypes.data_as(ip), y.ctypes.data_as(fp))
    assert rc == 0, f"devtail_eval: HIP error {rc}"
    return m, k, y


def test_device_tail_mass_and_quantile():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 40
    zs = np.concatenate([np.linspace(0.0, 6.0, 1201), np.linspace(6.0, 30.0, 481),
                         [1e-30, 2.0, 9.0, 26.5, 45.0]]).astype(np.float32)
    edges = np.log2(np.concatenate([4.8285e-4 * (1 + np.linspace(-1e-3, 1e-3, 21)),
                                    5.7988e-17 * (1 + np.linspace(-1e-3, 1e-3, 21))]))
    log2p = np.concatenate([np.linspace(-1.0, -40.0, 157), np.linspace(-40.0, -3000.0, 97), edges])
    tk = np.array([max(0, int(-lp) - 100) if lp < -110 else 0 for lp in log2p], np.int32)
    ts = np.array([np.float32(2.0 ** (lp + k)) for lp, k in zip(log2p, tk)], np.float32)
    n = max(len(zs), len(ts))
    zpad = np.resize(zs, n)
    tpad, kpad = np.resize(ts, n), np.resize(tk, n)
    print(f"devtail: {len(zs)} masses, {len(ts)} quantiles", flush=True)
    m, k, y = _eval(zpad, tpad, kpad)

    worst_m = 0.0
    for i, z in enumerate(zs):
        assert k[i] >= 0 and 0.02 <= m[i] <= 0.71, (z, m[i], k[i])
        tr = mpmath.erfc(mpmath.mpf(float(min(z, np.float32(30.0))))) / 2
        worst_m = max(worst_m, float(abs(mpmath.mpf(float(m[i])) * mpmath.mpf(2) ** (-int(k[i])) / tr - 1)))
    worst_q = 0.0
    for i in range(len(ts)):
        p = mpmath.mpf(float(ts[i])) * mpmath.mpf(2) ** (-int(tk[i]))
        tr = normal_quantile_mp(p, mpmath)
        worst_q = max(worst_q, float(abs(float(y[i]) - tr) / max(abs(tr), 1)))
    print(f"device tail mass worst rel {worst_m:.3e}, quantile worst {worst_q / np.finfo(np.float32).eps:.2f} eps",
          flush=True)
    assert worst_m < 5e-7, worst_m
    assert worst_q < 4 * np.finfo(np.float32).eps, worst_q
    # p = 1/2 gives 0 exactly
    assert _eval(np.zeros(1), np.full(1, 0.5), np.zeros(1))[2][0] == 0.0
