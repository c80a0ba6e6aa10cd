"""Pins of the CPU oracle (runs without a GPU).

The reference ships no numeric golden vectors for this path (its tests are
invariants with loose tolerances, test/runtests.jl), so each part of the
oracle is checked against something independent of its own formula: library
routines (mpmath, SciPy), closed forms, brute force at small sizes and the
reference's own invariants and published numbers.
"""
import math

import mpmath as mp
import numpy as np
import pytest
from scipy.special import ndtr

from dbgenzpaper_amd import workloads as W
from dbgenzpaper_amd.sobol import direction_integers, skip_count
from tests._helpers import random_spd, standardized

mp.mp.dps = 40


# --- special functions --------------------------------------------------------
def test_erfinv_against_mpmath(oracle):
    ys = list(np.linspace(-0.999999, 0.999999, 300)) + [1e-300, 1e-12, 0.5, 0.5000001, 1 - 1e-10,
                                                        1 - 2 ** -52, -1 + 2 ** -52, 1 - 1e-15]
    worst = 0.0
    for y in ys:
        t = mp.erfinv(mp.mpf(y))
        worst = max(worst, float(abs((oracle.erfinv(y) - t) / t)))
    assert worst < 3e-16


def test_as241_against_mpmath(oracle):
    with mp.workdps(400):  # 2p - 1 must keep the digits of tiny p
        for p in [1e-300, 1e-30, 1e-8, 0.01, 0.2, 0.49, 0.51, 0.7, 0.975, 1 - 1e-10]:
            t = mp.sqrt(2) * mp.erfinv(2 * mp.mpf(p) - 1)
            assert abs((oracle.as241(p) - t) / t) < 5e-16


# --- generators ----------------------------------------------------------------
def test_sqrt_primes(oracle):
    q = oracle.sqrt_primes(200)
    assert q[0] == 0.41421356237309515          # frac(sqrt(2))
    primes = [p for p in range(2, 2000) if all(p % f for f in range(2, int(p ** 0.5) + 1))][:200]
    assert np.array_equal(q, np.array([math.sqrt(p) - math.floor(math.sqrt(p)) for p in primes]))
    assert np.all((q > 0) & (q < 1))


def test_richtmyer_reference_invariants(oracle):
    """test/runtests.jl:91-113: column 1024 of a Float32 block has 1024
    distinct values at n_base 0, after next_points! (n_base = m) and at 2^20."""
    q = oracle.sqrt_primes(1024)
    for n0 in (0, 1024, 2 ** 20):
        X = oracle.richtmyer_block(1024, 1024, n0, q).astype(np.float32)
        assert len(np.unique(X[:, -1])) == 1024


def test_richtmyer_closed_form(oracle):
    """x_j = frac((n0 + j) q) with 1-based j: the first point of block 0 is q
    itself, and every point is within the Float64 product's rounding of the
    exact fractional part."""
    q = oracle.sqrt_primes(5)
    X = oracle.richtmyer_block(5, 64, 100, q)
    assert np.all(oracle.richtmyer_block(5, 1, 0, q)[0] == q)
    for i in range(5):
        for j in range(1, 65):
            exact = mp.frac(mp.mpf(100 + j) * mp.mpf(q[i]))
            assert abs(float(exact) - X[j - 1, i]) <= 2 ** -52 * (100 + j)


def test_sobol_matches_scipy(oracle):
    """Direction numbers (Joe-Kuo table) and Gray-code order against SciPy's
    unscrambled Sobol generator (point 0 of SciPy is the origin, which the
    reference's next! never returns)."""
    from scipy.stats import qmc

    d = 12
    V = direction_integers(d)
    X = oracle.sobol_block(V, 0, 1023)
    s = qmc.Sobol(d, scramble=False).random(1024)
    assert np.array_equal(X / 2 ** 32, s[1:])
    # the block after skip(2m - 1) continues the same sequence
    X2 = oracle.sobol_block(V, skip_count(2 * 256 - 1), 256)
    assert np.array_equal(X2 / 2 ** 32, s[512:768])


def test_skip_count():
    assert skip_count(2 * 1024 - 1) == 2047
    assert skip_count(0) == 0 and skip_count(1) == 1 and skip_count(5) == 3


# --- Cholesky --------------------------------------------------------------------
def test_genz_reference_invariant(oracle):
    """test/runtests.jl:15-36 (n = 512 here to keep the CPU suite short)."""
    n = 512
    C = random_spd(n, 101)
    F = oracle.cholesky_genz(C, -np.ones(n), np.ones(n))
    U = np.triu(F.U)
    assert F.rank == n
    assert np.abs(standardized(C)[np.ix_(F.perm, F.perm)] - U.T @ U).max() < 1e-10


def _genz_unblocked(C, a, b, bs2):
    """Textbook right-looking Genz-Bretz ordering (Genz & Bretz 2009), with
    the reference's window rule: pivots are searched, and conditional means
    collected, inside windows of bs2 columns only.  Independent of the
    oracle's four-level blocked structure."""
    n = C.shape[0]
    d = np.sqrt(np.diag(C))
    M = C / np.outer(d, d)
    a = a / d
    b = b / d
    perm = np.arange(n)
    s = np.zeros(n)
    U = np.zeros((n, n))
    from scipy.special import erf

    for w0 in range(0, n, bs2):
        w1 = min(n, w0 + bs2)
        for i in range(w0, w1):
            m = np.sqrt(np.maximum(np.diag(M)[i:w1], 0))
            order = erf((b[i:w1] - s[i:w1]) / m / np.sqrt(2)) - erf((a[i:w1] - s[i:w1]) / m / np.sqrt(2))
            im = i + int(np.argmin(order))
            dii = order[im - i]
            for arr in (perm, s, a, b):
                arr[[i, im]] = arr[[im, i]]
            M[[i, im], :] = M[[im, i], :]
            M[:, [i, im]] = M[:, [im, i]]
            U[:, [i, im]] = U[:, [im, i]]
            uii = np.sqrt(M[i, i])
            U[i, i] = uii
            U[i, i + 1:] = M[i, i + 1:] / uii
            an, bn = (a[i] - s[i]) / uii, (b[i] - s[i]) / uii
            y = (np.exp(-an ** 2 / 2) - np.exp(-bn ** 2 / 2)) * 2 / np.sqrt(2 * np.pi) / dii
            s[i + 1:w1] += y * U[i, i + 1:w1]
            M[i + 1:, i + 1:] -= np.outer(U[i, i + 1:], U[i, i + 1:])
    return U, perm, a, b


@pytest.mark.parametrize("n,bs,seed", [(12, (2, 4, 8, 8), 1), (37, (4, 8, 16, 32), 2),
                                       (64, (8, 64, 64, 64), 3), (90, (5, 20, 40, 80), 4)])
def test_genz_blocked_equals_unblocked(oracle, n, bs, seed):
    C = random_spd(n, seed, jitter=n / 4)
    rng = np.random.default_rng(seed)
    a, b = -rng.uniform(0.2, 2, n), rng.uniform(0.2, 2, n)
    F = oracle.cholesky_genz(C, a, b, *bs)
    U, perm, ap, bp = _genz_unblocked(C, a, b, bs[1])
    assert np.array_equal(F.perm, perm)
    assert np.abs(np.triu(F.U) - U).max() < 1e-12
    assert np.allclose(F.a, ap, rtol=1e-14) and np.allclose(F.b, bp, rtol=1e-14)


def test_genz_special_orders(oracle):
    n = 40
    rng = np.random.default_rng(5)
    b = rng.uniform(0.1, 2.0, n)
    a = -b * rng.uniform(0.5, 1.5, n)
    F = oracle.cholesky_genz(np.eye(n), a, b, 4, 16, 32, 64)
    mass = ndtr(b) - ndtr(a)
    exp = []
    for w0 in range(0, n, 16):
        exp += sorted(range(w0, min(n, w0 + 16)), key=lambda k: (mass[k], k))
    assert list(F.perm) == exp
    E = oracle.cholesky_genz(np.eye(n) + np.ones((n, n)), -np.ones(n), np.ones(n), 8, 32)
    assert list(E.perm) == list(range(n))


@pytest.mark.parametrize("bs", [8, 16, 32, 64])
def test_classic_reference_invariant(oracle, bs):
    """test/runtests.jl:38-59 and numpy's Cholesky of the standardized matrix."""
    n = 100
    C = random_spd(n, 102)
    F = oracle.cholesky_classic(C, -np.ones(n), np.ones(n), bs, bs, bs, bs)
    U = np.triu(F.U)
    Cs = standardized(C)
    assert np.abs(Cs - U.T @ U).max() < 1e-10
    assert np.abs(U - np.linalg.cholesky(Cs).T).max() < 1e-12


def test_rank_deficient(oracle):
    n = 30
    rng = np.random.default_rng(9)
    B = rng.standard_normal((n, 20))
    F = oracle.cholesky_genz(B @ B.T, -np.ones(n), np.ones(n))
    assert F.rank == 20


def test_rowmax_reference_invariant(oracle):
    """cholesky_rowmax! (src/cholesky.jl:266-432): C[p, p] = U'U for a full-rank
    matrix; a and b standardized but not permuted."""
    n = 300
    C = random_spd(n, 103)
    rng = np.random.default_rng(3)
    a, b = -rng.uniform(0.5, 2, n), rng.uniform(0.5, 2, n)
    F = oracle.cholesky_rowmax(C, a, b, 16, 64, 128, 256)
    U = np.triu(F.U)
    assert F.rank == n and sorted(F.perm) == list(range(n))
    assert np.abs(standardized(C)[np.ix_(F.perm, F.perm)] - U.T @ U).max() < 1e-10
    d = np.sqrt(np.diag(C))
    assert np.array_equal(F.a, a / d) or np.allclose(F.a, a / d, rtol=1e-15, atol=0)
    assert np.allclose(F.b, b / d, rtol=1e-15, atol=0) and np.array_equal(F.scale_vec, d)


@pytest.mark.parametrize("n,bs,seed", [(40, (8, 64, 64, 64), 11), (97, (16, 128, 128, 128), 12),
                                       (150, (64, 256, 256, 256), 13)])
def test_rowmax_one_window_equals_lapack_pstrf(oracle, n, bs, seed):
    """With one pivot window over the whole matrix the rule is plain diagonal
    pivoting, LAPACK's dpstrf (largest remaining diagonal, first on ties):
    the same permutation and factor, whatever the inner blocking."""
    from scipy.linalg import lapack

    C = random_spd(n, seed, jitter=n / 8)
    F = oracle.cholesky_rowmax(C, -np.ones(n), np.ones(n), *bs)
    Cs = standardized(C)
    np.fill_diagonal(Cs, 1.0)  # the reference's standardization writes exact ones (src/cholesky.jl:289)
    c, piv, rank, info = lapack.dpstrf(Cs, tol=-1.0, lower=0)
    assert info == 0 and rank == n == F.rank
    assert np.array_equal(F.perm, piv - 1)
    assert np.abs(np.triu(F.U) - np.triu(c)).max() < 1e-12


def _rowmax_unblocked(C, bs2):
    """Textbook right-looking diagonal pivoting with the reference's window rule
    (the pivot is searched inside windows of bs2 columns only); independent of
    the oracle's four-level blocked structure."""
    n = C.shape[0]
    d = np.sqrt(np.diag(C))
    M = C / np.outer(d, d)
    np.fill_diagonal(M, 1.0)
    perm = np.arange(n)
    U = np.zeros((n, n))
    for w0 in range(0, n, bs2):
        w1 = min(n, w0 + bs2)
        for i in range(w0, w1):
            im = i + int(np.argmax(np.diag(M)[i:w1]))
            perm[[i, im]] = perm[[im, i]]
            M[[i, im], :] = M[[im, i], :]
            M[:, [i, im]] = M[:, [im, i]]
            U[:, [i, im]] = U[:, [im, i]]
            uii = np.sqrt(M[i, i])
            U[i, i] = uii
            U[i, i + 1:] = M[i, i + 1:] / uii
            M[i + 1:, i + 1:] -= np.outer(U[i, i + 1:], U[i, i + 1:])
    return U, perm


@pytest.mark.parametrize("n,bs,seed", [(12, (2, 4, 8, 8), 21), (37, (4, 8, 16, 32), 22),
                                       (90, (5, 20, 40, 80), 23)])
def test_rowmax_blocked_equals_unblocked(oracle, n, bs, seed):
    C = random_spd(n, seed, jitter=n / 4)
    F = oracle.cholesky_rowmax(C, -np.ones(n), np.ones(n), *bs)
    U, perm = _rowmax_unblocked(C, bs[1])
    assert np.array_equal(F.perm, perm)
    assert np.abs(np.triu(F.U) - U).max() < 1e-12


def test_rowmax_exactly_singular(oracle):
    """Identical variables: after the first pivot every remaining Schur diagonal
    is exactly 0, and the rank check m_ii > 0 stops at rank 1 (ties keep the
    first index).  Two copies of each of 5 variables: the first 5 pivots take
    one of each pair; the copies' Schur diagonals are left at rounding level
    and, having no eps threshold, the reference's rule keeps factoring while
    they are positive (pivots of order 1e-8 at most)."""
    n = 16
    F = oracle.cholesky_rowmax(np.ones((n, n)), -np.ones(n), np.ones(n), 4, 16, 16, 16)
    assert F.rank == 1 and list(F.perm) == list(range(n))
    rng = np.random.default_rng(4)
    B = rng.standard_normal((5, 5))
    S = B @ B.T + 5 * np.eye(5)
    E = np.kron(S, np.ones((2, 2)))  # variables 2k and 2k+1 identical
    F = oracle.cholesky_rowmax(E, -np.ones(10), np.ones(10), 2, 16, 16, 16)
    assert 5 <= F.rank <= 10
    assert sorted(int(p) // 2 for p in F.perm[:5]) == list(range(5))
    assert np.all(np.diag(F.U)[:5] > 0.1) and np.all(np.abs(np.diag(F.U)[5:F.rank]) < 1e-6)


def _low_rank_case(n, rank, seed, zero_cols=(), neg=True):
    """A rank-deficient factor in the reference test's shape
    (test/runtests.jl:61-90): unit columns for the independent variables,
    dependent columns with random entries above row rank whose last nonzero
    row (the level) varies, some columns zero."""
    rng = np.random.default_rng(seed)
    M = np.zeros((n, n))
    M[np.arange(rank), np.arange(rank)] = 1.0
    for j in range(rank if rank > 0 else n, n):
        lvl = int(rng.integers(0, rank))
        M[:lvl + 1, j] = rng.uniform(0.1, 1.0, lvl + 1) * (rng.choice([-1.0, 1.0], lvl + 1) if neg else 1.0)
    for c in zero_cols:
        M[:, c] = 0.0
    a = -rng.uniform(0.5, 2.0, n)
    b = rng.uniform(0.5, 2.0, n)
    return M, a, b


def _level(col, rank, eps1):
    nz = np.nonzero(np.abs(col[:rank]) > eps1)[0]
    return int(nz[-1]) if len(nz) else -1


def test_adapt_low_rank_reference_test(oracle):
    """test/runtests.jl:61-90: n = 30, rank 20, column 25 (1-based) zero."""
    n, rank = 30, 20
    rng = np.random.default_rng(61)
    M = np.zeros((n, n))
    M[np.arange(rank), np.arange(rank)] = 1.0
    M[:rank, rank:] = rng.random((rank, n - rank))
    M[:, 24] = 0.0
    Mo, ao, bo, perm, kjs, k0 = oracle.adapt_low_rank(M, np.zeros(n), np.ones(n), rank, np.arange(n), 1e-10)
    assert k0 == 1 and perm[-1] == 24 and np.all(Mo[:, -1] == 0.0) and len(kjs) == rank


@pytest.mark.parametrize("n,rank,seed,zeros", [(30, 20, 1, (24,)), (64, 24, 2, (30, 40, 63)),
                                               (200, 120, 3, (130, 199)), (50, 0, 4, tuple(range(50)))])
def test_adapt_low_rank_invariants(oracle, n, rank, seed, zeros):
    """What adapt_low_rank! guarantees, checked on its result: every output
    column is an input column (perm), scaled by 1/m on rows up to its level m;
    the levels are sorted (zero columns last, k_0 of them); each dependent
    column's level entry is 1 and its bounds stay ordered (exchanged when m was
    negative); k_js[i] is the last column of level i."""
    eps1 = 1e-10
    M, a, b = _low_rank_case(n, rank, seed, zeros)
    Mo, ao, bo, perm, kjs, k0 = oracle.adapt_low_rank(M, a, b, rank, np.arange(n), eps1)
    assert sorted(perm) == list(range(n)) and k0 == len(zeros)
    lv = [_level(Mo[:, c], rank, eps1) for c in range(n)]
    active = n - k0
    assert all(v == -1 for v in lv[active:]) and all(v >= 0 for v in lv[:active]) or rank == 0
    assert lv[:active] == sorted(lv[:active])
    for c in range(n):
        p = int(perm[c])
        src = M[:, p]
        m = src[lv[c]] if lv[c] >= 0 and p >= rank else 1.0
        exp = src.copy()
        if p >= rank and lv[c] >= 0:
            exp[:lv[c] + 1] = src[:lv[c] + 1] * (1.0 / m)
            assert abs(Mo[lv[c], c] - 1.0) <= 2 ** -52
            lo, hi = sorted((a[p] * (1.0 / m), b[p] * (1.0 / m)))
            assert ao[c] == lo and bo[c] == hi
        else:
            assert ao[c] == a[p] and bo[c] == b[p]
        assert np.array_equal(Mo[:, c], exp)
    for i in range(rank):
        assert lv[kjs[i]] == i and (kjs[i] + 1 >= active or lv[kjs[i] + 1] > i)


# --- integration -------------------------------------------------------------------
def _run(oracle, Sigma, a, b, *, m, R, seed=0, phi="erf", kind="dense", gen="Richtmyer", **kw):
    n = Sigma.shape[0]
    rng = np.random.default_rng(seed)
    F = oracle.prepare_dense(Sigma, a, b) if kind == "dense" else oracle.prepare_sparse(Sigma, a, b)
    if gen == "Sobol":
        sh = rng.integers(0, 2 ** 64, size=(n - 1, R), dtype=np.uint64)
        kw["V"] = direction_integers(n - 1)
    else:
        sh = rng.random((n - 1, R))
    bs_i, bs_i2 = oracle.clamp_block_sizes(n, kw.pop("bs_i", 2 ** 6), kw.pop("bs_i2", 2 ** 5))
    return oracle.qmc_pnorm(F.U, F.a, F.b, m=m, n_reps=R, phi_form=phi, kind=kind, gen=gen,
                            shifts=sh, bs_i=bs_i, bs_i2=bs_i2, **kw), F


def test_identity_closed_form(oracle):
    n = 16
    rng = np.random.default_rng(3)
    a, b = -rng.uniform(0.2, 2, n), rng.uniform(0.2, 2, n)
    r, _ = _run(oracle, np.eye(n), a, b, m=4096, R=4)
    assert abs(r.value / np.prod(ndtr(b) - ndtr(a)) - 1) < 1e-13 and r.err3 < 1e-15


def test_bivariate_exact(oracle):
    """d = 2: one sampled dimension; against the bivariate normal CDF."""
    from scipy.stats import multivariate_normal

    S = np.array([[1.0, 0.6], [0.6, 1.0]])
    a, b = np.array([-np.inf, -np.inf]), np.array([0.3, -0.4])
    r, _ = _run(oracle, S, a, b, m=2 ** 14, R=8)
    ref = multivariate_normal([0, 0], S).cdf(b)
    assert abs(r.value - ref) < max(3 * r.err3, 1e-6)


@pytest.mark.parametrize("d", [16, 64, 256])
def test_equicorrelated_quadrature(oracle, d):
    """fixed_dense (Sigma = 1 1^T + I) against the 1-D quadrature truth of
    scripts/simulations/mwp.jl:80-116, inside the QMC error."""
    P = W.equicorrelated(d, m=2 ** 12, n_reps=12)
    truth = W.equicorrelated_truth(1.0, 1.0, P.a, P.b)
    r, _ = _run(oracle, P.Sigma, P.a, P.b, m=P.m, R=P.n_reps)
    assert abs(r.value - truth) < max(2 * r.err3, 1e-4)


def test_equicorrelated_published_medians():
    """The reference's published fixed_dense medians
    (data/sims/sparse_dense_vals.csv: 0.698106 at n = 256, 0.738887 at
    n = 512, m = 245760) agree with the quadrature truth to their spread."""
    for d, med in [(256, 0.698106405416119), (512, 0.7388865012733516)]:
        P = W.equicorrelated(d, m=16)
        assert abs(W.equicorrelated_truth(1.0, 1.0, P.a, P.b) - med) < 1e-3


def test_erf_vs_erfc_form(oracle):
    """Same points: |erf form - erfc form| <= 4e-14 (data/sims/table_precision.tex)."""
    P = W.dense_random(32, m=2048, n_reps=4, seed=7)
    r1, _ = _run(oracle, P.Sigma, P.a, P.b, m=P.m, R=P.n_reps, seed=1)
    r2, _ = _run(oracle, P.Sigma, P.a, P.b, m=P.m, R=P.n_reps, seed=1, phi="erfc")
    assert abs(r1.value - r2.value) < 4e-14


def test_dense_equals_sparse(oracle):
    P = W.banded(50, 4, m=1024, n_reps=3)
    F = oracle.prepare_sparse(P.Sigma, P.a, P.b)
    sh = np.random.default_rng(0).random((49, 3))
    rd = oracle.qmc_pnorm(F.U, F.a, F.b, m=1024, n_reps=3, shifts=sh, kind="dense", bs_i=16, bs_i2=4)
    rs = oracle.qmc_pnorm(F.U, F.a, F.b, m=1024, n_reps=3, shifts=sh, kind="sparse")
    assert abs(rd.value - rs.value) < 1e-14 * abs(rd.value)


def test_blocking_only_changes_rounding(oracle):
    P = W.dense_random(40, m=512, n_reps=2, seed=8)
    res = [_run(oracle, P.Sigma, P.a, P.b, m=512, R=2, seed=3, bs_i=bi, bs_i2=bi2, bs_j=bj)[0].value
           for bi, bi2, bj in [(64, 32, 64), (8, 2, 16), (16, 16, 512), (39, 5, 7)]]
    assert max(res) - min(res) < 1e-14 * abs(res[0])


def _brute_estimate(U, a, b, q, shifts, m, phi_erf=True):
    """The estimator of src/qmc.jl:315-444 for one shift, sample by sample, in
    40-digit mpmath arithmetic (no blocking, no Float64 rounding)."""
    n = U.shape[0]
    sq2 = mp.sqrt(2)
    tot = mp.mpf(0)
    for j in range(1, m + 1):
        za, zb = mp.mpf(a[0]) / (mp.mpf(U[0, 0]) * sq2), mp.mpf(b[0]) / (mp.mpf(U[0, 0]) * sq2)
        c = (1 + mp.erf(za)) / 2
        dc = (mp.erf(zb) - mp.erf(za)) / 2
        p = dc
        ys = []
        for i in range(1, n):
            x = mp.frac(mp.mpf(j) * mp.mpf(q[i - 1]))
            u = mp.mpf(shifts[i - 1]) + x
            u = u - 1 if u > 1 else u
            u = abs(2 * u - 1)
            u = c + u * dc
            ys.append(sq2 * mp.erfinv(2 * u - 1))
            s = sum(mp.mpf(U[l, i]) * ys[l] for l in range(i))
            inv = 1 / (mp.mpf(U[i, i]) * sq2)
            xa = max(min((mp.mpf(a[i]) - s) * inv, 9), -9)
            xb = max(min((mp.mpf(b[i]) - s) * inv, 9), -9)
            c = (1 + mp.erf(xa)) / 2
            dc = (1 + mp.erf(xb)) / 2 - c
            p *= dc
        tot += p
    return tot / m


def test_brute_force_tiny(oracle):
    """d = 4, 24 points: oracle equals the high-precision estimator."""
    P = W.dense_random(4, m=24, n_reps=1, seed=13, b_range=(-0.5, 1.5))
    F = oracle.prepare_dense(P.Sigma, P.a, P.b)
    sh = P.shifts
    q = oracle.sqrt_primes(3)
    r = oracle.qmc_pnorm(F.U, F.a, F.b, m=24, n_reps=1, shifts=sh, q=q, bs_i=3, bs_i2=2, bs_j=8)
    exact = _brute_estimate(np.triu(F.U), F.a, F.b, q, sh[:, 0], 24)
    assert abs(r.value - float(exact)) < 1e-14 * float(exact)


def test_adaptive_rounds(oracle):
    """max_abs_err = 0 runs max_pts / m rounds (test/runtests.jl:157-167);
    the running mean of src/qmc.jl:437-442 equals the plain mean of the
    per-round means."""
    P = W.dense_random(8, m=2 ** 9, n_reps=4, seed=42)
    F = oracle.prepare_dense(P.Sigma, P.a, P.b)
    sh = P.shifts
    r = oracle.qmc_pnorm(F.U, F.a, F.b, m=2 ** 9, n_reps=4, shifts=sh, max_pts=2 ** 11, max_abs_err=0.0,
                         bs_i=7, bs_i2=7)
    assert r.rounds == 4 and r.n_pts == 2 ** 11 and r.gen_state == 3
    per = [oracle.qmc_pnorm(F.U, F.a, F.b, m=2 ** 9, n_reps=4, shifts=sh, gen_state=g, bs_i=7,
                            bs_i2=7).rep_means for g in range(4)]
    assert np.allclose(r.rep_means, np.mean(per, axis=0), rtol=1e-14, atol=0)


def test_underflow_keeps_log(oracle):
    """A 400-dim lower orthant whose P is far below the Float64 range: the
    scaled product keeps log P, and it equals the sum of per-block logs on
    a block-diagonal problem (independent blocks multiply)."""
    Pb = W.factor_model(20, 4, seed=2)
    nb = 20
    S = np.kron(np.eye(nb), Pb.Sigma)
    a = np.full(20 * nb, -np.inf)
    b = np.tile(Pb.b, nb)
    r, _ = _run(oracle, S, a, b, m=2 ** 10, R=2, phi="erfc")
    assert r.value == 0.0 and np.isfinite(r.log_value) and r.log_value < -708
    one, _ = _run(oracle, Pb.Sigma, Pb.a, Pb.b, m=2 ** 14, R=4, phi="erfc")
    assert abs(r.log_value - nb * one.log_value) < 0.05 * abs(r.log_value)


def test_sobol_estimate(oracle):
    P = W.dense_random(10, m=1024, n_reps=4, seed=5)
    r, _ = _run(oracle, P.Sigma, P.a, P.b, m=1024, R=4, gen="Sobol", sobol_skip=skip_count(2047))
    rr, _ = _run(oracle, P.Sigma, P.a, P.b, m=2 ** 14, R=8)
    assert abs(r.value - rr.value) < 3 * (r.err3 + rr.err3) + 1e-6


# --- tail form (no reference counterpart) -------------------------------------------
def test_ndtri_tail_against_mpmath(oracle):
    with mp.workdps(400):
        for t in [1e-300, 1e-200, 1e-100, 1e-40, 1e-17, 1e-8, 1e-3, 0.1, 0.3, 0.4999]:
            truth = mp.sqrt(2) * mp.erfinv(2 * mp.mpf(t) - 1)
            assert abs((oracle.ndtri_tail(t) - truth) / truth) < 4e-16


def test_tail_form_equals_reference_form_without_underflow(oracle):
    """Where no interval reaches the [-9, 9] clamp, the tail form is the same
    estimator as the erf form (differences are rounding)."""
    P = W.dense_random(32, m=2048, n_reps=3, seed=7)
    r1, _ = _run(oracle, P.Sigma, P.a, P.b, m=P.m, R=P.n_reps, seed=1)
    r2, _ = _run(oracle, P.Sigma, P.a, P.b, m=P.m, R=P.n_reps, seed=1, phi="tail")
    assert abs(r1.value - r2.value) < 1e-13 * r1.value


def test_tail_form_identity_closed_form(oracle):
    n = 16
    rng = np.random.default_rng(3)
    a, b = -rng.uniform(0.2, 2, n), rng.uniform(0.2, 2, n)
    r, _ = _run(oracle, np.eye(n), a, b, m=1024, R=2, phi="tail")
    assert abs(r.value / np.prod(ndtr(b) - ndtr(a)) - 1) < 1e-13


def test_tail_form_deep_lower_orthant(oracle):
    """Factor-model lower orthant: every weight of the reference forms is 0
    (per-dimension masses fall under the clamp); the tail form gives a finite
    log P that agrees with itself across point sets within the QMC error."""
    P = W.factor_model(100, m=2 ** 10, n_reps=4)
    r_erf, _ = _run(oracle, P.Sigma, P.a, P.b, m=P.m, R=4, phi="erfc")
    assert r_erf.value == 0.0
    r1, _ = _run(oracle, P.Sigma, P.a, P.b, m=2 ** 10, R=4, seed=1, phi="tail")
    r2, _ = _run(oracle, P.Sigma, P.a, P.b, m=2 ** 12, R=4, seed=2, phi="tail")
    assert np.isfinite(r1.log_value) and np.isfinite(r2.log_value)
    assert abs(r1.log_value - r2.log_value) < 3 * (r1.rel_err3 + r2.rel_err3) + 0.05


# --- adjoint (no reference counterpart) ------------------------------------------
def test_grad_matches_central_differences(oracle):
    """d log P / d(a, b, U) of the oracle's reverse sweep against central
    differences of its own forward on the same points (h = 1e-5)."""
    rng = np.random.default_rng(21)
    n = 10
    Gm = rng.normal(size=(n, n))
    S = Gm @ Gm.T + n * np.eye(n)
    a, b = -rng.uniform(0.3, 2, n), rng.uniform(0.3, 2, n)
    a[3] = -np.inf
    F = oracle.prepare_dense(S, a, b)
    sh = rng.random((n - 1, 4))
    g = oracle.qmc_grad(F.U, F.a, F.b, m=1024, n_reps=4, shifts=sh)

    def lp(U, a_, b_):
        return oracle.qmc_pnorm(U, a_, b_, m=1024, n_reps=4, shifts=sh, phi_form="tail",
                                bs_i=n - 1, bs_i2=n - 1).log_value

    assert abs(g.log_value - lp(F.U, F.a, F.b)) < 1e-12 * abs(g.log_value)
    h = 1e-5
    for i in range(n):
        for which, vec, gv in (("b", F.b, g.gb), ("a", F.a, g.ga)):
            if np.isinf(vec[i]):
                assert gv[i] == 0.0
                continue
            vp, vm = vec.copy(), vec.copy()
            vp[i] += h
            vm[i] -= h
            fd = (lp(F.U, *((F.a, vp) if which == "b" else (vp, F.b))) -
                  lp(F.U, *((F.a, vm) if which == "b" else (vm, F.b)))) / (2 * h)
            assert abs(fd - gv[i]) < 1e-6 * max(1.0, abs(fd)), (which, i, fd, gv[i])
    for (l, t) in [(0, 1), (0, 9), (2, 5), (4, 4), (8, 9), (9, 9), (0, 0)]:
        Up, Um = F.U.copy(), F.U.copy()
        Up[l, t] += h
        Um[l, t] -= h
        fd = (lp(Up, F.a, F.b) - lp(Um, F.a, F.b)) / (2 * h)
        assert abs(fd - g.gU[l, t]) < 1e-6 * max(1.0, abs(fd)), (l, t, fd, g.gU[l, t])


def test_grad_sigma_chain_rule(oracle):
    """The Cholesky / standardization / bound chain rule used by the tests
    (tests._helpers.chain_to_sigma) against central differences in Sigma,
    a and b with the pivot order held fixed."""
    from tests._helpers import chain_to_sigma, logp_fixed_perm

    rng = np.random.default_rng(22)
    n = 8
    Gm = rng.normal(size=(n, n))
    S = Gm @ Gm.T + n * np.eye(n)
    a, b = -rng.uniform(0.5, 2, n), rng.uniform(0.5, 2, n)
    a[2] = -np.inf
    perm = oracle.prepare_dense(S, a, b).perm
    sh = rng.random((n - 1, 4))
    kw = dict(m=2048, n_reps=4, shifts=sh)
    lp0, U, ap, bp = logp_fixed_perm(oracle, S, a, b, perm, **kw)
    g = oracle.qmc_grad(U, ap, bp, m=2048, n_reps=4, shifts=sh)
    ga, gb, G = chain_to_sigma(U, perm, S, g.ga, g.gb, g.gU, ap, bp)
    h = 1e-5
    for (i, j) in [(0, 0), (1, 3), (2, 5), (7, 7), (4, 6)]:
        E = np.zeros((n, n))
        E[i, j] += h
        if i != j:
            E[j, i] += h
        fd = (logp_fixed_perm(oracle, S + E, a, b, perm, **kw)[0] -
              logp_fixed_perm(oracle, S - E, a, b, perm, **kw)[0]) / (2 * h)
        an = G[i, j] * (1 if i == j else 2)
        assert abs(fd - an) < 1e-6 * max(1.0, abs(fd)), (i, j, fd, an)
    for i in (0, 3, 6):
        e = np.zeros(n)
        e[i] = h
        fdb = (logp_fixed_perm(oracle, S, a, b + e, perm, **kw)[0] -
               logp_fixed_perm(oracle, S, a, b - e, perm, **kw)[0]) / (2 * h)
        fda = (logp_fixed_perm(oracle, S, a + e, b, perm, **kw)[0] -
               logp_fixed_perm(oracle, S, a - e, b, perm, **kw)[0]) / (2 * h)
        assert abs(fdb - gb[i]) < 1e-6 and abs(fda - ga[i]) < 1e-6
