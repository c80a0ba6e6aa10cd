"""Batched qmc_pnorm! over many independent problems of uneven dimension.

No counterpart in the reference (it solves one problem per
``qmc_pnorm!(QMCData(...))`` call); each problem here is exactly that call
with ``max_pts = m`` (one round), Richtmyer points and the erf-form CDF.  The
whole batch -- device factorizations, one persistent integration launch over
all (problem, shift, tile) work items and the per-problem reductions -- is a
single HIP graph replayed by :meth:`Batch.run`.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L


@dataclass
class BatchResult:
    value: np.ndarray      # P per problem (0 when it underflows Float64)
    log_value: np.ndarray
    err3: np.ndarray
    rel_err3: np.ndarray
    rank: np.ndarray


class Batch:
    """K problems (Sigma_k, a_k, b_k, shifts_k) integrated together."""

    def __init__(self, Sigmas, As, Bs, shifts, *, T=np.float32, m=2 ** 14, n_reps=4,
                 chol_block_size=2 ** 6, chol_block_size2=2 ** 9):
        K = len(Sigmas)
        dims = np.array([np.asarray(S).shape[0] for S in Sigmas], np.int32)
        sig = np.concatenate([np.asarray(S, np.float64).ravel(order="F") for S in Sigmas])
        a = np.concatenate([np.asarray(x, np.float64).ravel() for x in As])
        b = np.concatenate([np.asarray(x, np.float64).ravel() for x in Bs])
        sh = [np.asarray(x, np.float64).reshape(d - 1, n_reps).ravel(order="F") for x, d in zip(shifts, dims)]
        sh = np.concatenate(sh) if sh else np.zeros(1)
        if sh.size == 0:
            sh = np.zeros(1)
        o = L.Opts()
        T = np.dtype(T).type
        L.lib().dbg_opts_default(C.byref(o), L.PREC_MIXED if T == np.float32 else L.PREC_FP64)
        o.m, o.max_pts, o.n_reps = m, m, n_reps
        o.chol_block_size, o.chol_block_size2 = chol_block_size, chol_block_size2
        self.K, self.dims, self.m, self.n_reps, self.T = K, dims, m, n_reps, T
        self._keep = (sig, a, b, sh)
        h = C.c_void_p()
        L.check(L.lib().dbg_batch_create(C.byref(h), C.byref(o), K, L.ptr(dims, C.c_int32),
                                         L.ptr(sig, C.c_double), L.ptr(a, C.c_double), L.ptr(b, C.c_double),
                                         L.ptr(sh, C.c_double)))
        self._h = h
        self._out = np.zeros((K, 8), np.float64)

    def run(self) -> BatchResult:
        L.check(L.lib().dbg_batch_run(self._h, L.ptr(self._out, C.c_double)))
        o = self._out
        return BatchResult(o[:, 0].copy(), o[:, 1].copy(), o[:, 2].copy(), o[:, 3].copy(),
                           o[:, 7].astype(np.int64))

    def last_ms(self) -> float:
        ms = C.c_float()
        L.check(L.lib().dbg_batch_last_ms(self._h, C.byref(ms)))
        return float(ms.value)

    def tuning(self):
        w = C.c_int32()
        t = (C.c_float * 3)()
        L.check(L.lib().dbg_batch_tuning(self._h, C.byref(w), t))
        return int(w.value), list(t)

    def set_width(self, width: int) -> None:
        """Override the autotuner's dimension-block width (8, 16, or 32: for
        the mixed path the FP16 matrix-core contraction); the job's graph is
        captured again."""
        L.check(L.lib().dbg_batch_set_width(self._h, int(width)))

    def close(self):
        if getattr(self, "_h", None):
            L.lib().dbg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def qmc_pnorm_batched(problems, T=np.float32, **kw) -> BatchResult:
    """One run over a list of workloads.Problem (each with its own shifts)."""
    B = Batch([p.Sigma for p in problems], [p.a for p in problems], [p.b for p in problems],
              [p.shifts for p in problems], T=T, m=problems[0].m, n_reps=problems[0].n_reps, **kw)
    try:
        return B.run()
    finally:
        B.close()
