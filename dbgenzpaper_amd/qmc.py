"""Host-side mirror of the reference's Julia API for the Genz QMC path.

Same names, argument meaning and error behaviour as the reference
(src/qmc.jl, src/cholesky.jl, src/qmc_sparse.jl); every computation runs in
libdbgenz.so on the GPU.  Python spelling: ``qmc_pnorm`` for ``qmc_pnorm!``,
``cholesky_genz`` for ``cholesky_genz!`` (both still mutate their inputs like
the reference), the element type ``T`` is ``numpy.float64`` or
``numpy.float32``, permutations are 0-based, and ``ArgumentError`` is
``ValueError``.

``T = numpy.float32`` selects the mixed precision path: FP32 contraction and
special functions with the factor, the points and the reductions kept in FP64
(the reference factors in Float32 too; this build does not).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, replace

import numpy as np

from . import _lib as L
from .sobol import direction_integers, skip_count


# ---------------------------------------------------------------------------
# Options  (QMCOpts / QMC_opts, src/qmc.jl:15-27, 111-125)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class QMCOpts:
    chol_block_size: int
    chol_block_size2: int
    m: int
    block_size_i: int
    block_size_i2: int
    block_size_j: int
    n_blocks: int
    n_reps: int
    n_bits: int
    max_pts: int
    max_abs_err: float
    T: type = np.float64


def QMC_opts(T=np.float64, *, chol_block_size=2 ** 6, chol_block_size2=2 ** 9, m=10 ** 4,
             block_size_i=2 ** 6, block_size_j=2 ** 6, block_size_i2=2 ** 5, n_blocks=1,
             n_reps=12, n_bits=32, max_pts=10 ** 4, max_abs_err=10 ** (-6)) -> QMCOpts:
    """QMC_opts(T; ...)  src/qmc.jl:111-125 (same keywords and defaults;
    n_blocks defaults to 1 instead of Threads.nthreads())."""
    T = _check_T(T)
    return QMCOpts(int(chol_block_size), int(chol_block_size2), int(m), int(block_size_i),
                   int(block_size_i2), int(block_size_j), int(n_blocks), int(n_reps), int(n_bits),
                   int(max_pts), float(max_abs_err), T)


def _check_T(T):
    T = np.dtype(T).type
    if T not in (np.float64, np.float32):
        raise TypeError("T must be numpy.float64 or numpy.float32")
    return T


# ---------------------------------------------------------------------------
# Cholesky factorizations  (src/cholesky.jl)
# ---------------------------------------------------------------------------
@dataclass
class CholeskyGenz:
    """CholeskyGenz{T}  src/cholesky.jl:5-14 (perm 0-based)."""
    U: np.ndarray
    perm: np.ndarray
    rank: int
    a: np.ndarray
    b: np.ndarray
    block_size: int
    pivot_method: str
    scale_vec: np.ndarray


def _as_matrix(M):
    M = np.asarray(M)
    if M.ndim != 2 or M.shape[0] != M.shape[1]:
        raise ValueError("expected a square matrix")
    return M


def _write_back(dst, src):
    if isinstance(dst, np.ndarray) and dst.flags.writeable:
        dst[...] = src


def cholesky_genz(M, a, b, block_size=2 ** 7, block_size2=2 ** 9, block_size3=2 ** 10,
                  block_size4=2 ** 11) -> CholeskyGenz:
    """cholesky_genz!(M, a, b, bs1, bs2, bs3, bs4)  src/cholesky.jl:445-628.

    Genz-Bretz pivoted blocked Cholesky in FP64 on the GPU.  Like the
    reference it overwrites M (with U; strict lower triangle zeroed), a and b
    (standardized and permuted) when they are writable float64 arrays."""
    Mi = _as_matrix(M)
    n = Mi.shape[0]
    W = np.array(Mi, dtype=np.float64, order="F", copy=True)
    av = np.array(a, dtype=np.float64, copy=True).reshape(n)
    bv = np.array(b, dtype=np.float64, copy=True).reshape(n)
    perm = np.zeros(n, np.int32)
    scale = np.zeros(n, np.float64)
    rank = C.c_int32(0)
    L.check(L.lib().dbg_chol_genz_f64(L.ptr(W, C.c_double), n, n, L.ptr(av, C.c_double),
                                      L.ptr(bv, C.c_double), block_size, block_size2, block_size3,
                                      block_size4, L.ptr(perm, C.c_int32), L.ptr(scale, C.c_double),
                                      C.byref(rank)))
    _write_back(M, W)
    _write_back(a, av)
    _write_back(b, bv)
    return CholeskyGenz(W, perm.astype(np.int64), int(rank.value), av, bv, block_size, "genz", scale)


def cholesky_rowmax(M, a, b, block_size=2 ** 7, block_size2=2 ** 9, block_size3=2 ** 10,
                    block_size4=2 ** 11) -> CholeskyGenz:
    """cholesky_rowmax!(M, a, b, bs1, bs2, bs3, bs4)  src/cholesky.jl:266-432.

    The blocked loop nest of cholesky_genz! with diagonal pivoting (the largest
    remaining Schur diagonal of the pivot window, first index on ties; the
    factorization stops where it is not positive), in FP64 on the GPU.  Like
    the reference it overwrites M (with U; strict lower triangle zeroed), a and
    b (standardized, not permuted) when they are writable float64 arrays."""
    Mi = _as_matrix(M)
    n = Mi.shape[0]
    W = np.array(Mi, dtype=np.float64, order="F", copy=True)
    av = np.array(a, dtype=np.float64, copy=True).reshape(n)
    bv = np.array(b, dtype=np.float64, copy=True).reshape(n)
    perm = np.zeros(n, np.int32)
    scale = np.zeros(n, np.float64)
    rank = C.c_int32(0)
    L.check(L.lib().dbg_chol_rowmax_f64(L.ptr(W, C.c_double), n, n, L.ptr(av, C.c_double),
                                        L.ptr(bv, C.c_double), block_size, block_size2, block_size3,
                                        block_size4, L.ptr(perm, C.c_int32), L.ptr(scale, C.c_double),
                                        C.byref(rank)))
    _write_back(M, W)
    _write_back(a, av)
    _write_back(b, bv)
    return CholeskyGenz(W, perm.astype(np.int64), int(rank.value), av, bv, block_size, "row_max", scale)


def adapt_low_rank(M, a, b, rank, perm, eps1):
    """adapt_low_rank!(M, a, b, rank, perm, eps1)  src/cholesky.jl:20-123, on the GPU.

    Reorders the dependent columns rank .. n-1 of a rank-deficient factor:
    each is scaled so that its last entry above row rank with |value| > eps1
    is 1 (a, b scaled too, exchanged when that entry was negative) and moved
    to the end of the block of that level; zero columns go to the end.  M, a,
    b and perm (0-based) are updated in place when they are writable arrays of
    the right type (float64, float64, float64, integer).  Returns (k_js, k_0):
    each level's last column, 0-based (the reference's k_js minus one), and the
    number of zero columns."""
    Mi = _as_matrix(M)
    n = Mi.shape[0]
    if not 0 <= rank <= n:
        raise ValueError("rank must be in [0, n]")
    W = np.array(Mi, dtype=np.float64, order="F", copy=True)
    av = np.array(a, dtype=np.float64, copy=True).reshape(n)
    bv = np.array(b, dtype=np.float64, copy=True).reshape(n)
    pv = np.array(perm, dtype=np.int32, copy=True).reshape(n)
    kjs = np.zeros(max(rank, 1), np.int32)
    k0 = C.c_int32(0)
    L.check(L.lib().dbg_adapt_low_rank_f64(L.ptr(W, C.c_double), n, n, L.ptr(av, C.c_double),
                                           L.ptr(bv, C.c_double), int(rank), L.ptr(pv, C.c_int32), float(eps1),
                                           L.ptr(kjs, C.c_int32), C.byref(k0)))
    _write_back(M, W)
    _write_back(a, av)
    _write_back(b, bv)
    _write_back(perm, pv)
    return kjs[:rank].astype(np.int64), int(k0.value)


def covariance(kind, n, *, jitter=1.0, n_sp=1, nu=1.5, dist=1.0, sigma2=1.0, B=None, psi=None,
               standardize=False) -> np.ndarray:
    """The covariance generators of the reference's simulation scripts
    (scripts/simulations/simulations_functions.jl), built on the GPU and
    returned as a column-major float64 array:
      "fixed_dense"   ones(n, n) + jitter I               (:21-24)
      "fixed_sparse"  fixed_sparse(n, n_sp)               (:33-41)
      "matern"        mattern_cov(n; nu, dist, sigma2)    (:56-79), nu in {0.5, 1.5, 2.5}
      "lowrank"       B B' + diag(psi): the factor model; rand_spd(n; jitter) (:13-18)
                      with B = A (n x n, drawn by the caller) and psi = jitter * n
    standardize rescales to a correlation matrix.  (The C entry point
    dbg_cov_build can also leave the matrix in device memory for
    dbg_plan_create.)"""
    kinds = {"fixed_dense": L.COV_FIXED_DENSE, "fixed_sparse": L.COV_FIXED_SPARSE, "matern": L.COV_MATERN,
             "lowrank": L.COV_LOWRANK}
    if kind not in kinds:
        raise ValueError(f"unknown covariance kind {kind!r}")
    S = np.zeros((n, n), np.float64, order="F")
    params = np.array({"fixed_dense": [jitter, 0, 0], "fixed_sparse": [n_sp, 0, 0], "matern": [nu, dist, sigma2],
                       "lowrank": [0, 0, 0]}[kind], np.float64)
    k, Bp, pp = 0, None, None
    if kind == "lowrank":
        Bm = np.array(B, np.float64, order="F", copy=True)
        if Bm.ndim != 2 or Bm.shape[0] != n:
            raise ValueError("B must be n x k")
        pv = np.array(psi, np.float64, copy=True).reshape(n)
        k, Bp, pp = Bm.shape[1], Bm.ctypes.data, pv.ctypes.data
    L.check(L.lib().dbg_cov_build(kinds[kind], n, L.ptr(params, C.c_double), k, Bp, pp, int(bool(standardize)),
                                  S.ctypes.data, n, 0))
    return S


def cholesky_classic(M, a, b, block_size=2 ** 7, block_size2=2 ** 9, block_size3=2 ** 10,
                     block_size4=2 ** 11) -> CholeskyGenz:
    """cholesky_classic!(M, a, b, bs1, bs2, bs3, bs4)  src/cholesky.jl:144-242 (GPU, FP64)."""
    Mi = _as_matrix(M)
    n = Mi.shape[0]
    W = np.array(Mi, dtype=np.float64, order="F", copy=True)
    av = np.array(a, dtype=np.float64, copy=True).reshape(n)
    bv = np.array(b, dtype=np.float64, copy=True).reshape(n)
    scale = np.zeros(n, np.float64)
    L.check(L.lib().dbg_chol_classic_f64(L.ptr(W, C.c_double), n, n, L.ptr(av, C.c_double),
                                         L.ptr(bv, C.c_double), block_size, block_size2,
                                         block_size3, block_size4, L.ptr(scale, C.c_double)))
    _write_back(M, W)
    _write_back(a, av)
    _write_back(b, bv)
    return CholeskyGenz(W, np.arange(n), n, av, bv, block_size, "none", scale)


# ---------------------------------------------------------------------------
# QMC data objects  (QMCData, QMCDataSparse, QMCData_Distributions)
# ---------------------------------------------------------------------------
@dataclass
class QMCGrad:
    """d log P of the tail-form estimate (see _QMCBase.grad)."""
    log_value: float
    a: np.ndarray          # d log P / d a (original order)
    b: np.ndarray          # d log P / d b
    Sigma: np.ndarray      # d log P / d Sigma, symmetric
    a_std: np.ndarray      # d log P / d a' (standardized, permuted)
    b_std: np.ndarray      # d log P / d b'
    U: np.ndarray          # d log P / d U (upper triangle)


class _QMCBase:
    """Owns one device plan (factor, packed layouts, points, workspace)."""

    _kind = L.KIND_DENSE
    _phi = L.PHI_ERF

    def __init__(self, C_mat, a, b, opts: QMCOpts | None, rng, qmc_type, *, shifts=None,
                 sobol_skip=None, chol_sizes=None, phi_form=None, contraction=None, refine_tiles=None,
                 _clamp=True):
        Cm = _as_matrix(C_mat)
        n = Cm.shape[0]
        if opts is None:
            opts = QMC_opts(np.float64)
        if _clamp:
            opts = _clamped_opts(opts, n, self._kind)
        self.qmc_opts = opts
        self.T = opts.T
        self.n = n
        if qmc_type not in ("Richtmyer", "Sobol"):
            raise ValueError("qmc_type must be 'Richtmyer' or 'Sobol'")
        self.qmc_type = qmc_type
        self._Sigma = np.array(Cm, dtype=np.float64, order="F", copy=True)
        self._a0 = np.array(a, dtype=np.float64, copy=True).reshape(n)
        self._b0 = np.array(b, dtype=np.float64, copy=True).reshape(n)
        self._sobol_skip = sobol_skip
        cs = chol_sizes or (opts.chol_block_size, opts.chol_block_size2, 2 ** 10, 2 ** 11)
        o = L.Opts()
        L.lib().dbg_opts_default(C.byref(o), L.PREC_MIXED if self.T == np.float32 else L.PREC_FP64)
        o.chol_block_size, o.chol_block_size2, o.chol_block_size3, o.chol_block_size4 = cs
        o.m = opts.m
        o.block_size_i, o.block_size_i2, o.block_size_j = (opts.block_size_i, opts.block_size_i2,
                                                          opts.block_size_j)
        o.n_blocks = opts.n_blocks
        o.n_reps = opts.n_reps
        o.n_bits = opts.n_bits
        o.max_pts = opts.max_pts
        o.max_abs_err = opts.max_abs_err
        o.generator = L.GEN_SOBOL if qmc_type == "Sobol" else L.GEN_RICHTMYER
        if phi_form is not None:
            self._phi = {"erf": L.PHI_ERF, "erfc": L.PHI_ERFC, "tail": L.PHI_TAIL}[phi_form]
        o.phi_form = self._phi
        o.kind = self._kind
        o.sobol_skip = -1 if sobol_skip is None else int(sobol_skip)
        # arithmetic of the mixed dense contraction: None/"auto", "fp32" or "split16"
        o.contraction = {None: L.CONTRACT_AUTO, "auto": L.CONTRACT_AUTO, "fp32": L.CONTRACT_FP32,
                         "split16": L.CONTRACT_SPLIT16}[contraction]
        # mixed dense plans: tiles per shift evaluated again in FP64 (None: the library's default)
        o.refine_tiles = -1 if refine_tiles is None else int(refine_tiles)
        self._opts_c = o
        plan = C.c_void_p()
        rc = L.lib().dbg_plan_create(C.byref(plan), C.byref(o), n, L.ptr(self._Sigma, C.c_double), n,
                                     L.ptr(self._a0, C.c_double), L.ptr(self._b0, C.c_double))
        if rc == L.DBG_ERR_RANK:
            raise ValueError("QMCData requires a full-rank Cholesky factorization.")
        L.check(rc)
        self._plan = plan
        if o.generator == L.GEN_SOBOL:
            V = np.ascontiguousarray(direction_integers(max(n - 1, 0)), np.uint32)
            L.check(L.lib().dbg_plan_set_sobol_directions(plan, L.ptr(V, C.c_uint32)))
        if shifts is None:
            shifts = draw_shifts(rng, n, opts.n_reps, qmc_type, self.T)
        self.set_shifts(shifts)
        self.last_result = None
        self._C = None

    # -- state -------------------------------------------------------------
    def set_shifts(self, shifts):
        n, R = self.n, self.qmc_opts.n_reps
        if self.qmc_type == "Sobol" or self._kind == L.KIND_SPARSEQMC:
            sh = np.asfortranarray(np.asarray(shifts, dtype=np.uint64).reshape(n - 1, R))
            L.check(L.lib().dbg_plan_set_shifts(self._plan, sh.ctypes.data_as(C.c_void_p)))
        else:
            sh = np.asfortranarray(np.asarray(shifts, dtype=np.float64).reshape(n - 1, R))
            L.check(L.lib().dbg_plan_set_shifts(self._plan, sh.ctypes.data_as(C.c_void_p)))
        self.shifts = sh

    @property
    def gen_state(self) -> int:
        s = C.c_int64()
        L.check(L.lib().dbg_plan_get_state(self._plan, C.byref(s)))
        return s.value

    @gen_state.setter
    def gen_state(self, v: int):
        L.check(L.lib().dbg_plan_set_state(self._plan, int(v)))

    @property
    def C(self) -> CholeskyGenz:
        """The factor held on the device (downloaded on first access)."""
        if self._C is None:
            n = self.n
            U = np.zeros((n, n), np.float64, order="F")
            perm = np.zeros(n, np.int32)
            a = np.zeros(n)
            b = np.zeros(n)
            scale = np.zeros(n)
            rank = C.c_int32()
            L.check(L.lib().dbg_plan_get_factor(self._plan, L.ptr(U, C.c_double), L.ptr(perm, C.c_int32),
                                                L.ptr(a, C.c_double), L.ptr(b, C.c_double),
                                                L.ptr(scale, C.c_double), C.byref(rank)))
            method = ("none" if self._kind != L.KIND_DENSE
                      else ("genz" if int(rank.value) == n and not self._lowrank_route() else "row_max"))
            self._C = CholeskyGenz(U, perm.astype(np.int64), int(rank.value), a, b,
                                   self.qmc_opts.chol_block_size, method, scale)
        return self._C

    def _lowrank_route(self) -> bool:
        r, kjs, _ = self.lowrank_info()
        return r < self.n or not np.array_equal(kjs, np.arange(r))

    @property
    def a(self):
        return self.C.a

    @property
    def b(self):
        return self.C.b

    @property
    def qmc_reps(self) -> np.ndarray:
        """Per-shift running means (Float64; 0 where they underflow)."""
        m, e = self.rep_scaled()
        return np.ldexp(m, e)

    def rep_scaled(self):
        R = self.qmc_opts.n_reps
        m = np.zeros(R, np.float64)
        e = np.zeros(R, np.int32)
        L.check(L.lib().dbg_plan_get_reps(self._plan, L.ptr(m, C.c_double), L.ptr(e, C.c_int32)))
        return m, e.astype(np.int64)

    # -- adjoint -------------------------------------------------------------
    def grad(self) -> "QMCGrad":
        """Gradient of log P of one round of the tail-form estimate (the
        plan's current block of m points and its shifts), computed on the
        device: with respect to the constructor's a, b and Sigma (symmetric,
        pivot order fixed; d log P = sum_ij G_ij E_ij for a symmetric
        perturbation E) and to the standardized, permuted a', b' and the
        factor U.  No reference counterpart."""
        n = self.n
        lv = C.c_double()
        ga, gb = np.zeros(n), np.zeros(n)
        gs = np.zeros((n, n), order="F")
        ga_s, gb_s = np.zeros(n), np.zeros(n)
        gU = np.zeros((n, n), order="F")
        L.check(L.lib().dbg_qmc_grad(self._plan, C.byref(lv), L.ptr(ga, C.c_double), L.ptr(gb, C.c_double),
                                     L.ptr(gs, C.c_double), L.ptr(ga_s, C.c_double), L.ptr(gb_s, C.c_double),
                                     L.ptr(gU, C.c_double)))
        return QMCGrad(float(lv.value), ga, gb, gs, ga_s, gb_s, gU)

    def lowrank_info(self):
        """(rank, last column of each level (0-based), zero-variance columns)
        of the plan's factor: a Sigma of rank r < n takes the rank-deficient
        route (row-max factor, adapt_low_rank!, FP64 level loop) where the
        reference's QMCData throws; (n, 0 .. n-1, 0) at full rank."""
        r, k0 = C.c_int32(), C.c_int32()
        kjs = np.zeros(self.n, np.int32)
        L.check(L.lib().dbg_plan_lowrank_info(self._plan, C.byref(r), L.ptr(kjs, C.c_int32), C.byref(k0)))
        return int(r.value), kjs[:r.value].astype(np.int64), int(k0.value)

    def set_grad_budget(self, record_bytes: int = 0) -> None:
        """Device memory for the gradient's y and s records per chunk of
        samples (0: the default, 4 GiB); larger rounds run in chunks."""
        L.check(L.lib().dbg_plan_set_grad_budget(self._plan, int(record_bytes)))

    def grad_info(self):
        """(samples per chunk, chunks, device bytes of all gradient buffers)
        of the last grad(); zeros before the first."""
        cs, nc, nb = C.c_int64(), C.c_int32(), C.c_int64()
        L.check(L.lib().dbg_plan_grad_info(self._plan, C.byref(cs), C.byref(nc), C.byref(nb)))
        return int(cs.value), int(nc.value), int(nb.value)

    # -- hot path ------------------------------------------------------------
    def loop(self, n_pts0: int = 0) -> None:
        """One qmc_loop! over the current block, queued without waiting."""
        L.check(L.lib().dbg_qmc_loop(self._plan, int(n_pts0)))

    def refactor(self) -> None:
        """Redo the factorization from the device-resident Sigma, a, b."""
        L.check(L.lib().dbg_plan_refactor(self._plan))
        self._C = None

    def sync(self) -> None:
        L.check(L.lib().dbg_plan_sync(self._plan))

    def last_kernel_ms(self) -> float:
        ms = C.c_float()
        L.check(L.lib().dbg_plan_last_kernel_ms(self._plan, C.byref(ms)))
        return float(ms.value)

    def refine_info(self):
        """(tiles per shift, largest share of a shift's sum left in FP32,
        device ms) of the last round's FP64 re-evaluation (mixed dense
        plans; zeros otherwise)."""
        k, f, ms = C.c_int32(), C.c_double(), C.c_float()
        L.check(L.lib().dbg_plan_refine_info(self._plan, C.byref(k), C.byref(f), C.byref(ms)))
        return int(k.value), float(f.value), float(ms.value)

    def summary(self):
        r = L.Result()
        L.check(L.lib().dbg_plan_summary(self._plan, C.byref(r)))
        return r

    def close(self):
        if getattr(self, "_plan", None):
            L.lib().dbg_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _clamped_opts(opts: QMCOpts, n: int, kind: int) -> QMCOpts:
    """Block-size clamps of the constructors (src/qmc.jl:166-174, 239-242)."""
    if kind == L.KIND_DENSE and n - 1 < opts.block_size_i:
        bs_i2 = n - 1 if n - 1 < opts.block_size_i2 else opts.block_size_i2
        return replace(opts, block_size_i=n - 1, block_size_i2=bs_i2)
    if kind == L.KIND_SPARSE and n - 1 < opts.block_size_i:
        return replace(opts, block_size_i=n)
    return opts


def draw_shifts(rng, n: int, n_reps: int, qmc_type: str, T=np.float64) -> np.ndarray:
    """Random shifts: rand(rng, T, n-1, n_reps) (src/qmc_generators.jl:52) or
    rand(rng, UInt64, n-1, n_reps) (:63), from a numpy Generator."""
    if rng is None:
        rng = np.random.default_rng()
    elif isinstance(rng, (int, np.integer)):
        rng = np.random.default_rng(int(rng))
    shape = (max(n - 1, 0), n_reps)
    if qmc_type == "Sobol":
        return np.asfortranarray(rng.integers(0, 2 ** 64, size=shape, dtype=np.uint64))
    if T == np.float32:
        return np.asfortranarray(rng.random(shape, dtype=np.float32).astype(np.float64))
    return np.asfortranarray(rng.random(shape))


class QMCData(_QMCBase):
    """QMCData(C, a, b, opts=QMC_opts(T), rng=default_rng(), qmc_type="Richtmyer")
    src/qmc.jl:149-202.  Pivoted Genz factor, dense loop, erf-form CDF.

    Keyword ``phi_form="tail"`` (no reference counterpart) selects the tail
    form for T = Float64: normal tail masses without the [-9, 9] clamp and
    Phi^{-1} from the nearer tail, which keeps problems whose per-dimension
    masses fall below 1e-37 finite.  T = Float32 always uses it."""
    _kind = L.KIND_DENSE
    _phi = L.PHI_ERF

    def __init__(self, C_mat, a, b, opts=None, rng=None, qmc_type="Richtmyer", **kw):
        super().__init__(C_mat, a, b, opts, rng, qmc_type, **kw)


class QMCData_Distributions(_QMCBase):
    """QMCData_Distributions(D)  src/qmc.jl:210-215: the same problem with the
    normal CDF in erfc form (src/qmc.jl:579-593, 803-805); same shifts and
    generator state as D."""
    _kind = L.KIND_DENSE
    _phi = L.PHI_ERFC

    def __init__(self, D: QMCData):
        super().__init__(D._Sigma, D._a0, D._b0, D.qmc_opts, None, D.qmc_type, shifts=D.shifts,
                         sobol_skip=D._sobol_skip, phi_form="erfc", _clamp=False)
        self.gen_state = D.gen_state


class QMCDataSparse(_QMCBase):
    """QMCDataSparse(C, a, b, opts=QMC_opts(T), rng=default_rng(), qmc_type="Sobol")
    src/qmc.jl:225-282.  Unpivoted factor, CSC loop (src/qmc.jl:618-718)."""
    _kind = L.KIND_SPARSE
    _phi = L.PHI_ERF

    def __init__(self, C_mat, a, b, opts=None, rng=None, qmc_type="Sobol", **kw):
        super().__init__(C_mat, a, b, opts, rng, qmc_type, **kw)


def qmc_pnorm(D: _QMCBase, use_AppleBLAS=None):
    """qmc_pnorm!(D)  src/qmc.jl:783-850 -> (value, err, n_pts).

    ``use_AppleBLAS`` is accepted and ignored.  The log of the value, which
    stays finite when P underflows Float64, is in ``D.last_result``."""
    r = L.Result()
    L.check(L.lib().dbg_qmc_pnorm(D._plan, C.byref(r)))
    D.last_result = r
    return float(r.value), float(r.err3), int(r.n_pts)


# ---------------------------------------------------------------------------
# SparseQMC submodule  (src/qmc_sparse.jl)
# ---------------------------------------------------------------------------
class _SparseQMCModule:
    """SparseQMC module (src/qmc_sparse.jl): Sobol points without the tent
    map, unpivoted factor with the default block sizes, shifts re-drawn from
    the rng at each qmc_pnorm!."""

    @dataclass(frozen=True)
    class QMCOpts:
        m: int
        block_size_i: int
        block_size_j: int
        n_blocks: int
        n_reps: int
        n_bits: int
        max_pts: int
        max_abs_err: float
        T: type = np.float64

    @staticmethod
    def QMC_opts(T=np.float64, *, m=10 ** 4, block_size_i=2 ** 6, block_size_j=2 ** 6, n_blocks=1,
                 n_reps=12, n_bits=32, max_pts=10 ** 4, max_abs_err=10 ** (-6)):
        """SparseQMC.QMC_opts  src/qmc_sparse.jl:42-53"""
        return _SparseQMCModule.QMCOpts(int(m), int(block_size_i), int(block_size_j), int(n_blocks),
                                        int(n_reps), int(n_bits), int(max_pts), float(max_abs_err),
                                        _check_T(T))

    class QMCDataSparse(_QMCBase):
        """SparseQMC.QMCDataSparse(C, a, b; opts)  src/qmc_sparse.jl:58-90"""
        _kind = L.KIND_SPARSEQMC
        _phi = L.PHI_ERF

        def __init__(self, C_mat, a, b, *, opts=None, sobol_skip=None):
            if opts is None:
                opts = _SparseQMCModule.QMC_opts(np.float64)
            full = QMCOpts(2 ** 7, 2 ** 9, opts.m, opts.block_size_i, 2 ** 5, opts.block_size_j,
                           opts.n_blocks, opts.n_reps, opts.n_bits, opts.max_pts, opts.max_abs_err,
                           opts.T)
            n = _as_matrix(C_mat).shape[0]
            zero = np.zeros((max(n - 1, 0), opts.n_reps), np.uint64)
            super().__init__(C_mat, a, b, full, None, "Sobol", shifts=zero, sobol_skip=sobol_skip,
                             chol_sizes=(2 ** 7, 2 ** 9, 2 ** 10, 2 ** 11), _clamp=False)

    @staticmethod
    def qmc_pnorm(D, rng=None, *, shifts=None):
        """SparseQMC.qmc_pnorm!(D, rng)  src/qmc_sparse.jl:272-317: draws new
        shifts (rand!(rng, D.shifts)) then integrates."""
        if shifts is None:
            shifts = draw_shifts(rng, D.n, D.qmc_opts.n_reps, "Sobol")
        D.set_shifts(shifts)
        return qmc_pnorm(D)


SparseQMC = _SparseQMCModule()

__all__ = ["QMCOpts", "QMC_opts", "CholeskyGenz", "cholesky_genz", "cholesky_classic", "QMCData",
           "QMCData_Distributions", "QMCDataSparse", "qmc_pnorm", "SparseQMC", "draw_shifts",
           "skip_count"]
