"""ctypes front end of the CPU FP64 oracle (``genz_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` to compute expected values.  The
product (``dbgenzpaper_amd``) never imports it.

Each wrapper names the reference routine it restates (paths relative to the
reference repository root).  Matrices are column-major (Fortran order) float64
like Julia's ``Matrix{Float64}``; permutations are 0-based.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_genz.so")
_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp, ip, i64p, u32p, u64p = (C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_uint64))
        L.orc_erfinv.restype = C.c_double
        L.orc_erfinv.argtypes = [C.c_double]
        L.orc_as241.restype = C.c_double
        L.orc_as241.argtypes = [C.c_double]
        L.orc_ndtri_tail.restype = C.c_double
        L.orc_ndtri_tail.argtypes = [C.c_double]
        L.orc_sqrt_primes.argtypes = [C.c_int, dp]
        L.orc_richtmyer_block.argtypes = [C.c_int, C.c_int64, C.c_int64, dp, dp]
        L.orc_sobol_block.argtypes = [C.c_int, u32p, C.c_uint64, C.c_int64, u64p]
        L.orc_cholesky_genz.argtypes = [dp, C.c_int, dp, dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                        ip, dp, ip, dp]
        L.orc_cholesky_rowmax.argtypes = [dp, C.c_int, dp, dp, C.c_int, C.c_int, C.c_int, C.c_int, ip, dp, ip,
                                          C.c_double]
        L.orc_adapt_low_rank.argtypes = [dp, C.c_int, dp, dp, C.c_int, ip, C.c_double, ip, ip]
        L.orc_cholesky_classic.argtypes = [dp, C.c_int, dp, dp, C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.orc_truncate_U.argtypes = [dp, C.c_int, C.c_double]
        L.orc_qmc_pnorm.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, i64p, C.POINTER(C.c_int32), dp,
                                    dp, dp, u32p, u64p, C.c_uint64, C.c_int64, C.c_double, i64p, dp,
                                    ip, C.c_void_p]
        L.orc_qmc_pnorm_lowrank.argtypes = [C.c_void_p, dp, dp, dp, dp, ip, C.c_int, dp, dp, u32p, u64p,
                                            C.c_uint64, C.c_int64, C.c_double, i64p, dp, ip, C.c_void_p]
        L.orc_qmc_grad_dense.argtypes = [C.c_void_p, dp, dp, dp, dp, dp, u32p, u64p, C.c_uint64, dp, dp,
                                         dp, dp]
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


class _Desc(C.Structure):
    _fields_ = [("n", C.c_int), ("m", C.c_int64), ("n_reps", C.c_int), ("gen", C.c_int),
                ("phi_form", C.c_int), ("bs_i", C.c_int), ("bs_i2", C.c_int), ("bs_j", C.c_int),
                ("n_bits", C.c_int), ("threads", C.c_int)]


class _Result(C.Structure):
    _fields_ = [("value", C.c_double), ("log_value", C.c_double), ("err3", C.c_double),
                ("rel_err3", C.c_double), ("n_pts", C.c_int64), ("rounds", C.c_int)]


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def erfinv(y: float) -> float:
    return lib().orc_erfinv(float(y))


def as241(p: float) -> float:
    return lib().orc_as241(float(p))


def ndtri_tail(t: float) -> float:
    """Phi^{-1}(t) for 0 < t <= 1/2 (the tail form's quantile)."""
    return lib().orc_ndtri_tail(float(t))


def sqrt_primes(n: int) -> np.ndarray:
    """get_sqrt_primes!  src/qmc_generators.jl:300-310"""
    q = np.zeros(n, np.float64)
    lib().orc_sqrt_primes(n, _ptr(q, C.c_double))
    return q


def richtmyer_block(n_dims: int, m: int, n0: int, q: np.ndarray) -> np.ndarray:
    """richtmyer_mat!  src/qmc_generators.jl:314-325 -> m x n_dims (Float64)."""
    q = np.ascontiguousarray(q, np.float64)
    X = np.zeros((m, n_dims), np.float64, order="F")
    lib().orc_richtmyer_block(n_dims, m, n0, _ptr(q, C.c_double), _ptr(X, C.c_double))
    return X


def sobol_block(V: np.ndarray, first_idx: int, m: int) -> np.ndarray:
    """Sobol points first_idx+1 .. first_idx+m as 32-bit integers (m x n_dims)."""
    V = np.ascontiguousarray(V, np.uint32)
    n_dims = V.shape[0]
    X = np.zeros((m, n_dims), np.uint64, order="F")
    lib().orc_sobol_block(n_dims, _ptr(V, C.c_uint32), first_idx, m, _ptr(X, C.c_uint64))
    return X


@dataclass
class CholeskyResult:
    U: np.ndarray          # n x n, column-major (upper triangle + diagonal meaningful)
    perm: np.ndarray       # 0-based
    rank: int
    a: np.ndarray
    b: np.ndarray
    scale_vec: np.ndarray
    gaps: np.ndarray | None = None


def cholesky_genz(M, a, b, bs1=2 ** 7, bs2=2 ** 9, bs3=2 ** 10, bs4=2 ** 11) -> CholeskyResult:
    """cholesky_genz!  src/cholesky.jl:445-628 (on copies)."""
    M = np.array(M, np.float64, order="F", copy=True)
    a = np.array(a, np.float64, copy=True)
    b = np.array(b, np.float64, copy=True)
    n = M.shape[0]
    perm = np.zeros(n, np.int32)
    scale = np.zeros(n, np.float64)
    rank = C.c_int(0)
    gaps = np.zeros(n, np.float64)
    rc = lib().orc_cholesky_genz(_ptr(M, C.c_double), n, _ptr(a, C.c_double), _ptr(b, C.c_double),
                                 bs1, bs2, bs3, bs4, _ptr(perm, C.c_int), _ptr(scale, C.c_double),
                                 C.byref(rank), _ptr(gaps, C.c_double))
    assert rc == 0
    return CholeskyResult(M, perm.astype(np.int64), rank.value, a, b, scale, gaps)


def cholesky_rowmax(M, a, b, bs1=2 ** 7, bs2=2 ** 9, bs3=2 ** 10, bs4=2 ** 11, tol=0.0) -> CholeskyResult:
    """cholesky_rowmax!  src/cholesky.jl:266-432 (on copies).  tol: the rank
    check is m_ii > tol (0: the reference's)."""
    M = np.array(M, np.float64, order="F", copy=True)
    a = np.array(a, np.float64, copy=True)
    b = np.array(b, np.float64, copy=True)
    n = M.shape[0]
    perm = np.zeros(n, np.int32)
    scale = np.zeros(n, np.float64)
    rank = C.c_int(0)
    rc = lib().orc_cholesky_rowmax(_ptr(M, C.c_double), n, _ptr(a, C.c_double), _ptr(b, C.c_double),
                                   bs1, bs2, bs3, bs4, _ptr(perm, C.c_int), _ptr(scale, C.c_double),
                                   C.byref(rank), float(tol))
    assert rc == 0
    return CholeskyResult(M, perm.astype(np.int64), rank.value, a, b, scale)


def adapt_low_rank(M, a, b, rank, perm, eps1):
    """adapt_low_rank!  src/cholesky.jl:20-123 (on copies; perm and k_js 0-based).
    Returns (M, a, b, perm, k_js, k_0)."""
    M = np.array(M, np.float64, order="F", copy=True)
    a = np.array(a, np.float64, copy=True)
    b = np.array(b, np.float64, copy=True)
    perm = np.array(perm, np.int32, copy=True)
    n = M.shape[0]
    kjs = np.zeros(max(rank, 1), np.int32)
    k0 = C.c_int(0)
    rc = lib().orc_adapt_low_rank(_ptr(M, C.c_double), n, _ptr(a, C.c_double), _ptr(b, C.c_double), rank,
                                  _ptr(perm, C.c_int), eps1, _ptr(kjs, C.c_int), C.byref(k0))
    assert rc == 0
    return M, a, b, perm.astype(np.int64), kjs[:rank].astype(np.int64), k0.value


def cholesky_classic(M, a, b, bs1=2 ** 7, bs2=2 ** 9, bs3=2 ** 10, bs4=2 ** 11) -> CholeskyResult:
    """cholesky_classic!  src/cholesky.jl:144-242 (on copies)."""
    M = np.array(M, np.float64, order="F", copy=True)
    a = np.array(a, np.float64, copy=True)
    b = np.array(b, np.float64, copy=True)
    n = M.shape[0]
    scale = np.zeros(n, np.float64)
    lib().orc_cholesky_classic(_ptr(M, C.c_double), n, _ptr(a, C.c_double), _ptr(b, C.c_double),
                               bs1, bs2, bs3, bs4, _ptr(scale, C.c_double))
    return CholeskyResult(M, np.arange(n), n, a, b, scale)


def truncate_U(U: np.ndarray, eps_T: float = np.finfo(np.float64).eps) -> np.ndarray:
    """truncate_matrixU!  src/qmc.jl:133-142 (in place, returns U)."""
    assert U.flags.f_contiguous and U.dtype == np.float64
    lib().orc_truncate_U(_ptr(U, C.c_double), U.shape[0], eps_T)
    return U


def csc_strict_upper(U: np.ndarray):
    """sparse(triu(U, 1)) of src/qmc.jl:272 as 0-based CSC (rows ascending)."""
    n = U.shape[0]
    colptr = np.zeros(n + 1, np.int64)
    rows, vals = [], []
    for j in range(n):
        col = U[:j, j]
        nz = np.nonzero(col)[0]
        rows.append(nz.astype(np.int32))
        vals.append(col[nz])
        colptr[j + 1] = colptr[j] + nz.size
    rowval = np.concatenate(rows) if rows else np.zeros(0, np.int32)
    nzval = np.concatenate(vals) if vals else np.zeros(0)
    return colptr, np.ascontiguousarray(rowval, np.int32), np.ascontiguousarray(nzval, np.float64)


@dataclass
class PnormResult:
    value: float
    log_value: float
    err3: float
    rel_err3: float
    n_pts: int
    rounds: int
    rep_means: np.ndarray      # shift means (float64, 0 where they underflow)
    rep_mant: np.ndarray
    rep_exp: np.ndarray
    gen_state: int


def qmc_pnorm(U, a, b, *, m, n_reps, gen="Richtmyer", phi_form="erf", kind="dense",
              bs_i=2 ** 6, bs_i2=2 ** 5, bs_j=2 ** 6, n_bits=32, max_pts=None, max_abs_err=1e-6,
              q=None, shifts=None, V=None, sobol_skip=None, gen_state=None, threads=0) -> PnormResult:
    """qmc_pnorm!  src/qmc.jl:783-850 over a prepared factor.

    kind: "dense" (QMCData, qmc.jl:315-444; phi_form "erfc" = QMCData_Distributions
    :485-608; phi_form "tail" = the complementary tail form without the [-9, 9]
    clamp, which has no reference counterpart), "sparse" (QMCDataSparse, :618-718) or "sparseqmc" (SparseQMC module,
    src/qmc_sparse.jl:159-317, Sobol only).  U must already be truncated; a and b
    are the permuted, standardized bounds.  Block sizes are used as given (the
    QMCData constructor clamps them first, src/qmc.jl:166-174; see
    ``clamp_block_sizes``).
    """
    U = np.asfortranarray(U, np.float64)
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    n = U.shape[0]
    kind_i = {"dense": 0, "sparse": 1, "sparseqmc": 2}[kind]
    gen_i = {"Richtmyer": 0, "Sobol": 1}[gen]
    if kind_i == 2:
        gen_i = 1
    desc = _Desc(n, m, n_reps, gen_i, {"erf": 0, "erfc": 1, "tail": 2}[phi_form], max(bs_i, 1), max(bs_i2, 1),
                 bs_j, n_bits, threads)
    if max_pts is None:
        max_pts = m
    q, shifts_f, shifts_u, V = _point_args(n, n_reps, gen_i, q, shifts, V)
    if sobol_skip is None:
        sobol_skip = sobol_skip_count(2 * m - 1)
    if kind_i != 0:
        colptr, rowval, nzval = csc_strict_upper(U)
    else:
        colptr, rowval, nzval = np.zeros(1, np.int64), np.zeros(1, np.int32), np.zeros(1)
    state = C.c_int64(gen_state if gen_state is not None else (1 if kind_i == 2 else 0))
    rep_m = np.zeros(n_reps, np.float64)
    rep_e = np.zeros(n_reps, np.int32)
    res = _Result()
    rc = lib().orc_qmc_pnorm(C.byref(desc), kind_i, _ptr(U, C.c_double), _ptr(a, C.c_double),
                             _ptr(b, C.c_double), _ptr(colptr, C.c_int64), _ptr(rowval, C.c_int32),
                             _ptr(nzval, C.c_double), _ptr(q, C.c_double), _ptr(shifts_f, C.c_double),
                             _ptr(V, C.c_uint32), _ptr(shifts_u, C.c_uint64), sobol_skip, max_pts,
                             max_abs_err, C.byref(state), _ptr(rep_m, C.c_double),
                             _ptr(rep_e, C.c_int), C.byref(res))
    assert rc == 0
    return PnormResult(res.value, res.log_value, res.err3, res.rel_err3, res.n_pts, res.rounds,
                       np.ldexp(rep_m, rep_e), rep_m, rep_e.astype(np.int64), state.value)


def _point_args(n, n_reps, gen_i, q, shifts, V):
    """Richtmyer sqrt-prime fractions and shifts ((n-1) x n_reps), or Sobol
    direction integers and digital shifts, as the C drivers take them."""
    nd = max(n - 1, 1)
    q = np.ascontiguousarray(sqrt_primes(nd) if q is None else q, np.float64)
    if gen_i == 0:
        shifts_f = np.asfortranarray(shifts, np.float64)
        assert shifts_f.shape == (n - 1, n_reps) or n == 1
        shifts_u = np.zeros(1, np.uint64)
        V = np.zeros((1, 32), np.uint32)
    else:
        shifts_u = np.asfortranarray(shifts, np.uint64)
        shifts_f = np.zeros(1, np.float64)
        V = np.ascontiguousarray(V, np.uint32)
        assert V.shape[0] >= n - 1
    return q, shifts_f, shifts_u, V


@dataclass
class LowRankFactor:
    """The level structure of a rank-deficient factor (no reference
    counterpart; see prepare_lowrank)."""
    U: np.ndarray          # n x n column-major: rows 0 .. rank-1 meaningful
    a: np.ndarray          # bounds of the columns (standardized, scaled by adapt_low_rank!)
    b: np.ndarray
    perm: np.ndarray       # original variable of each column, 0-based
    rank: int
    beg: np.ndarray        # rank + 1 column-block starts; beg[rank] = n - k0
    k0: int                # zero-variance columns at the end
    inv: np.ndarray        # 1 / (U[i, c] sqrt 2) for column c of block i
    scale_vec: np.ndarray


def prepare_lowrank(Sigma, a, b, *, chol_block_size=2 ** 6, chol_block_size2=2 ** 9, tol=None, eps1=None):
    """The rank-deficient route of the QMCData constructor (no reference
    counterpart: src/qmc.jl:195-197 throws).  cholesky_rowmax!
    (src/cholesky.jl:266-432) with the rank check m_ii > tol (default
    n * eps, LAPACK dpstrf's); rows 0 .. r-1 of the dependent columns r .. n-1
    completed from the standardized, permuted Sigma, U_r^T X = C_p[0:r, r:n];
    the standardized bounds permuted; adapt_low_rank! (src/cholesky.jl:20-123)
    with eps1 (default sqrt(eps)); the column blocks and 1/(U[i, c] sqrt 2)."""
    from scipy.linalg import solve_triangular

    S = np.array(Sigma, np.float64, order="F")
    n = S.shape[0]
    eps = np.finfo(np.float64).eps
    tol = n * eps if tol is None else tol
    eps1 = np.sqrt(eps) if eps1 is None else eps1
    F = cholesky_rowmax(S, a, b, chol_block_size, chol_block_size2, tol=tol)
    r = F.rank
    if r == n:
        raise ValueError("the factor has full rank")
    p = F.perm
    s = F.scale_vec
    U = np.zeros((n, n), order="F")
    U[:r, :r] = np.triu(F.U[:r, :r])
    Cp = S[np.ix_(p[:r], p[r:])] * ((1.0 / s[p[r:]])[None, :] / s[p[:r]][:, None])
    U[:r, r:] = solve_triangular(U[:r, :r], Cp, trans="T", lower=False)
    U, aa, bb, perm, kjs, k0 = adapt_low_rank(U, F.a[p], F.b[p], r, p, eps1)
    beg = np.zeros(r + 1, np.int32)
    beg[1:] = kjs + 1
    assert beg[r] == n - k0
    inv = np.zeros(n)
    for i in range(r):
        for c in range(beg[i], beg[i + 1]):
            inv[c] = 1.0 / (U[i, c] * np.sqrt(2.0))
    return LowRankFactor(U, aa, bb, perm, r, beg, k0, inv, s)


def qmc_pnorm_lowrank(LF: LowRankFactor, *, m, n_reps, gen="Richtmyer", phi_form="erf", bs_j=2 ** 6, n_bits=32,
                      max_pts=None, max_abs_err=1e-6, q=None, shifts=None, V=None, sobol_skip=None,
                      gen_state=None, threads=0) -> PnormResult:
    """qmc_pnorm! (src/qmc.jl:783-850) over the level structure of a
    rank-deficient factor: per sample, the interval of latent variable i is the
    intersection over its column block, and rank - 1 dimensions are sampled
    (the first rank - 1 rows of the (n-1) x n_reps shifts are used)."""
    U = np.asfortranarray(LF.U, np.float64)
    a = np.ascontiguousarray(LF.a, np.float64)
    b = np.ascontiguousarray(LF.b, np.float64)
    inv = np.ascontiguousarray(LF.inv, np.float64)
    beg = np.ascontiguousarray(LF.beg, np.int32)
    n = U.shape[0]
    gen_i = {"Richtmyer": 0, "Sobol": 1}[gen]
    desc = _Desc(n, m, n_reps, gen_i, {"erf": 0, "erfc": 1, "tail": 2}[phi_form], 1, 1, bs_j, n_bits, threads)
    if max_pts is None:
        max_pts = m
    q, shifts_f, shifts_u, V = _point_args(n, n_reps, gen_i, q, shifts, V)
    if sobol_skip is None:
        sobol_skip = sobol_skip_count(2 * m - 1)
    state = C.c_int64(0 if gen_state is None else gen_state)
    rep_m = np.zeros(n_reps, np.float64)
    rep_e = np.zeros(n_reps, np.int32)
    res = _Result()
    rc = lib().orc_qmc_pnorm_lowrank(C.byref(desc), _ptr(U, C.c_double), _ptr(a, C.c_double), _ptr(b, C.c_double),
                                     _ptr(inv, C.c_double), _ptr(beg, C.c_int), LF.rank, _ptr(q, C.c_double),
                                     _ptr(shifts_f, C.c_double), _ptr(V, C.c_uint32), _ptr(shifts_u, C.c_uint64),
                                     sobol_skip, max_pts, max_abs_err, C.byref(state), _ptr(rep_m, C.c_double),
                                     _ptr(rep_e, C.c_int), C.byref(res))
    assert rc == 0
    return PnormResult(res.value, res.log_value, res.err3, res.rel_err3, res.n_pts, res.rounds,
                       np.ldexp(rep_m, rep_e), rep_m, rep_e.astype(np.int64), state.value)


@dataclass
class GradResult:
    ga: np.ndarray        # d log P / d a (standardized, permuted bounds)
    gb: np.ndarray        # d log P / d b
    gU: np.ndarray        # d log P / d U (upper triangle incl. diagonal, column-major n x n)
    log_value: float      # log P of the same estimate


def qmc_grad(U, a, b, *, m, n_reps, gen="Richtmyer", shifts=None, q=None, V=None, sobol_skip=None,
             n_bits=32, threads=0) -> GradResult:
    """Gradient of log P (tail form, dense loop, one block of m points per
    shift) with respect to a, b and U, in FP64.  No reference counterpart:
    the reverse sweep of the forward loop, documented in genz_oracle.c."""
    U = np.asfortranarray(U, np.float64)
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    n = U.shape[0]
    nd = max(n - 1, 1)
    gen_i = {"Richtmyer": 0, "Sobol": 1}[gen]
    desc = _Desc(n, m, n_reps, gen_i, 2, 1, 1, 1, n_bits, threads)
    q = np.ascontiguousarray(sqrt_primes(nd) if q is None else q, np.float64)
    if gen_i == 0:
        shifts_f = np.asfortranarray(shifts, np.float64)
        shifts_u = np.zeros(1, np.uint64)
        V = np.zeros((1, 32), np.uint32)
    else:
        shifts_u = np.asfortranarray(shifts, np.uint64)
        shifts_f = np.zeros(1, np.float64)
        V = np.ascontiguousarray(V, np.uint32)
    if sobol_skip is None:
        sobol_skip = sobol_skip_count(2 * m - 1)
    ga, gb = np.zeros(n), np.zeros(n)
    gU = np.zeros((n, n), order="F")
    lp = C.c_double(0.0)
    rc = lib().orc_qmc_grad_dense(C.byref(desc), _ptr(U, C.c_double), _ptr(a, C.c_double), _ptr(b, C.c_double),
                                  _ptr(q, C.c_double), _ptr(shifts_f, C.c_double), _ptr(V, C.c_uint32),
                                  _ptr(shifts_u, C.c_uint64), sobol_skip, _ptr(ga, C.c_double),
                                  _ptr(gb, C.c_double), _ptr(gU, C.c_double), C.byref(lp))
    assert rc == 0
    return GradResult(ga, gb, gU, lp.value)


def sobol_skip_count(n: int) -> int:
    """Number of next! calls of ``skip(SobolSeq(d), n)`` as assumed here:
    the largest 2^k - 1 <= n (Sobol.jl's default, non-exact skip).  For
    n = 2m - 1 with m a power of two this is n itself."""
    if n <= 0:
        return 0
    k = (n + 1).bit_length() - 1
    return (1 << k) - 1


def clamp_block_sizes(n: int, bs_i: int, bs_i2: int):
    """QMCData constructor clamp  src/qmc.jl:166-174."""
    if n - 1 < bs_i:
        bs_i = n - 1
        if n - 1 < bs_i2:
            bs_i2 = n - 1
    return bs_i, bs_i2


def prepare_dense(Sigma, a, b, *, chol_block_size=2 ** 6, chol_block_size2=2 ** 9):
    """QMCData constructor (src/qmc.jl:149-202) factor part: cholesky_genz!
    on copies, full-rank check, truncate_matrixU!."""
    F = cholesky_genz(Sigma, a, b, chol_block_size, chol_block_size2)
    if F.rank != np.asarray(Sigma).shape[0]:
        raise ValueError("QMCData requires a full-rank Cholesky factorization.")
    truncate_U(F.U)
    return F


def prepare_sparse(Sigma, a, b, *, chol_block_size=2 ** 6, chol_block_size2=2 ** 9):
    """QMCDataSparse constructor (src/qmc.jl:225-282) factor part."""
    F = cholesky_classic(Sigma, a, b, chol_block_size, chol_block_size2)
    truncate_U(F.U)
    return F


def num_threads() -> int:
    return lib().orc_num_threads()
