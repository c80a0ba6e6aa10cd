"""Device Cholesky factorizations against the oracle and the reference's own
invariants (test/runtests.jl:15-59)."""
import numpy as np
import pytest

from tests._helpers import random_spd, standardized

pytestmark = pytest.mark.gpu


def _genz_pair(dbg, oracle, C, a, b, bs):
    Fo = oracle.cholesky_genz(C, a, b, *bs)
    Fg = dbg.cholesky_genz(C.copy(order="F"), a.copy(), b.copy(), *bs)
    return Fo, Fg


@pytest.mark.parametrize("n,bs,seed", [
    (16, (2 ** 7, 2 ** 9, 2 ** 10, 2 ** 11), 1),
    (100, (8, 32, 64, 128), 2),           # several windows, inner blocks, iii/iv nesting
    (300, (16, 64, 128, 256), 3),
    (1024, (2 ** 7, 2 ** 9, 2 ** 10, 2 ** 11), 101),  # the reference test's shape
])
def test_genz_matches_oracle(dbg, oracle, n, bs, seed):
    C = random_spd(n, seed)
    rng = np.random.default_rng(seed + 7)
    a = -rng.uniform(0.5, 2.0, n)
    b = rng.uniform(0.5, 2.0, n)
    print(f"genz n={n} bs={bs} seed={seed}", flush=True)
    Fo, Fg = _genz_pair(dbg, oracle, C, a, b, bs)
    assert Fg.rank == n == Fo.rank
    if not np.array_equal(Fg.perm, Fo.perm):
        k = int(np.argmax(Fg.perm != Fo.perm))
        pytest.fail(f"perm differs first at step {k}; oracle best/second gap {Fo.gaps[k]:.3e}")
    Uo, Ug = np.triu(Fo.U), np.triu(Fg.U)
    assert np.abs(Ug - Uo).max() < 1e-12
    assert np.allclose(Fg.a, Fo.a, rtol=1e-14, atol=0) and np.allclose(Fg.b, Fo.b, rtol=1e-14, atol=0)
    assert np.allclose(Fg.scale_vec, Fo.scale_vec, rtol=0, atol=0)
    Cs = standardized(C)
    p = Fg.perm
    assert np.abs(Cs[np.ix_(p, p)] - Ug.T @ Ug).max() < 1e-10   # test/runtests.jl:35


def test_genz_reference_invariant_ones_bounds(dbg):
    """test/runtests.jl:15-36 with a = -1, b = 1."""
    n = 1024
    C = random_spd(n, 101)
    F = dbg.cholesky_genz(C.copy(order="F"), -np.ones(n), np.ones(n))
    U = np.triu(F.U)
    assert np.abs(standardized(C)[np.ix_(F.perm, F.perm)] - U.T @ U).max() < 1e-10


def test_genz_identity_is_windowed_selection_sort(dbg):
    n = 40
    rng = np.random.default_rng(5)
    b = rng.uniform(0.1, 2.0, n)
    a = -b * rng.uniform(0.5, 1.5, n)
    F = dbg.cholesky_genz(np.eye(n), a.copy(), b.copy(), 4, 16, 32, 64)
    from scipy.special import ndtr

    mass = ndtr(b) - ndtr(a)
    exp = []
    for w0 in range(0, n, 16):
        exp += sorted(range(w0, min(n, w0 + 16)), key=lambda k: (mass[k], k))
    assert list(F.perm) == exp


def test_genz_exchangeable_identity_perm(dbg):
    n = 64
    C = np.eye(n) + np.ones((n, n))
    F = dbg.cholesky_genz(C, -np.ones(n), np.ones(n), 8, 32)
    assert list(F.perm) == list(range(n))


def test_genz_rank_deficient(dbg, oracle):
    n = 30
    rng = np.random.default_rng(9)
    B = rng.standard_normal((n, 20))
    C = B @ B.T
    Fo = oracle.cholesky_genz(C, -np.ones(n), np.ones(n))
    Fg = dbg.cholesky_genz(C.copy(order="F"), -np.ones(n), np.ones(n))
    assert Fg.rank == Fo.rank < n


def test_qmcdata_rank_deficient_route(dbg):
    """Where the reference's QMCData throws (src/qmc.jl:195-197), the plan
    takes the rank-deficient route: ten identical variables are one, and
    P(-1 <= X <= 1) = Phi(1) - Phi(-1) with nothing sampled."""
    from scipy.stats import norm

    n = 10
    C = np.ones((n, n))
    D = dbg.QMCData(C, -np.ones(n), np.ones(n), dbg.QMC_opts(m=256, max_pts=256), 0)
    assert D.lowrank_info()[0] == 1 and D.C.rank == 1
    v, e, _ = dbg.qmc_pnorm(D)
    D.close()
    assert abs(v - (norm.cdf(1.0) - norm.cdf(-1.0))) < 1e-15 and e < 1e-15


@pytest.mark.parametrize("bs", [8, 16, 32, 64])
def test_classic_matches_oracle(dbg, oracle, bs):
    """test/runtests.jl:38-59 (n = 100, equal block sizes at all levels)."""
    n = 100
    C = random_spd(n, 102)
    a, b = -np.ones(n), np.ones(n)
    Fo = oracle.cholesky_classic(C, a, b, bs, bs, bs, bs)
    Fg = dbg.cholesky_classic(C.copy(order="F"), a.copy(), b.copy(), bs, bs, bs, bs)
    Ug = np.triu(Fg.U)
    assert np.abs(standardized(C) - Ug.T @ Ug).max() < 1e-10
    assert np.abs(Ug - np.triu(Fo.U)).max() < 1e-12


@pytest.mark.parametrize("n,bs,seed", [
    (16, (2 ** 7, 2 ** 9, 2 ** 10, 2 ** 11), 31),
    (100, (8, 32, 64, 128), 32),            # several windows, inner blocks, iii/iv nesting
    (300, (16, 64, 128, 256), 33),
    (700, (64, 512, 1024, 2048), 34),       # 512-column windows: the register-window kernel
    (1024, (64, 1024, 1024, 2048), 35),     # a 1024-column window: the LDS-state kernel first
])
def test_rowmax_matches_oracle(dbg, oracle, n, bs, seed):
    """cholesky_rowmax! (src/cholesky.jl:266-432) on the device against the
    oracle: the same permutation, factor to 1e-12, a and b standardized and
    not permuted, the reference's invariant C[p, p] = U'U."""
    C = random_spd(n, seed)
    rng = np.random.default_rng(seed)
    a = -rng.uniform(0.5, 2.0, n)
    b = rng.uniform(0.5, 2.0, n)
    print(f"rowmax n={n} bs={bs} seed={seed}", flush=True)
    Fo = oracle.cholesky_rowmax(C, a, b, *bs)
    Fg = dbg.cholesky_rowmax(C.copy(order="F"), a.copy(), b.copy(), *bs)
    assert Fg.rank == n == Fo.rank
    assert np.array_equal(Fg.perm, Fo.perm)
    Ug = np.triu(Fg.U)
    assert np.abs(Ug - np.triu(Fo.U)).max() < 1e-12
    assert np.array_equal(Fg.a, Fo.a) and np.array_equal(Fg.b, Fo.b)
    assert np.array_equal(Fg.scale_vec, Fo.scale_vec)
    assert np.abs(standardized(C)[np.ix_(Fg.perm, Fg.perm)] - Ug.T @ Ug).max() < 1e-10


def test_rowmax_rank_stop(dbg, oracle):
    """Identical variables: every Schur diagonal after the first pivot is 0
    exactly and the factorization stops at rank 1 with the identity order.
    Two copies of each of 5 variables: the 5 leading pivots equal the
    oracle's, one of each pair; the copies' Schur diagonals are then at
    rounding level (1e-16), where the order the rule picks depends on the
    summation order and is not compared."""
    n = 16
    F = dbg.cholesky_rowmax(np.ones((n, n)), -np.ones(n), np.ones(n), 4, 16, 16, 16)
    assert F.rank == 1 and list(F.perm) == list(range(n))
    rng = np.random.default_rng(4)
    B = rng.standard_normal((5, 5))
    E = np.kron(B @ B.T + 5 * np.eye(5), np.ones((2, 2)))
    Fo = oracle.cholesky_rowmax(E, -np.ones(10), np.ones(10), 2, 16, 16, 16)
    Fg = dbg.cholesky_rowmax(E.copy(order="F"), -np.ones(10), np.ones(10), 2, 16, 16, 16)
    assert 5 <= Fg.rank <= 10 and 5 <= Fo.rank <= 10
    assert np.array_equal(Fg.perm[:5], Fo.perm[:5])
    assert sorted(int(p) // 2 for p in Fg.perm[:5]) == list(range(5))
    assert np.abs(np.triu(Fg.U)[:5, :5] - np.triu(Fo.U)[:5, :5]).max() < 1e-12


@pytest.mark.parametrize("n,rank,seed,zeros", [(30, 20, 1, (24,)), (64, 24, 2, (30, 40, 63)),
                                               (200, 120, 3, (130, 199)), (700, 300, 4, (305, 350, 699)),
                                               (40, 40, 5, ()), (16, 0, 6, tuple(range(16)))])
def test_adapt_low_rank_matches_oracle(dbg, oracle, n, rank, seed, zeros):
    """adapt_low_rank! (src/cholesky.jl:20-123) on the device against the
    oracle: the same column moves and scalings, bit for bit (the factor, a, b,
    perm, k_js and k_0)."""
    from tests.test_oracle import _low_rank_case

    print(f"adapt_low_rank n={n} rank={rank} seed={seed}", flush=True)
    M, a, b = _low_rank_case(n, rank, seed, zeros)
    Mo, ao, bo, po, ko, k0o = oracle.adapt_low_rank(M, a, b, rank, np.arange(n), 1e-10)
    Mg, ag, bg, pg = M.copy(order="F"), a.copy(), b.copy(), np.arange(n, dtype=np.int32)
    kg, k0g = dbg.adapt_low_rank(Mg, ag, bg, rank, pg, 1e-10)
    assert k0g == k0o and np.array_equal(kg, ko) and np.array_equal(pg, po)
    assert np.array_equal(Mg, Mo) and np.array_equal(ag, ao) and np.array_equal(bg, bo)


def test_adapt_low_rank_reference_test(dbg):
    """test/runtests.jl:61-90 on the device."""
    n, rank = 30, 20
    rng = np.random.default_rng(61)
    M = np.zeros((n, n), order="F")
    M[np.arange(rank), np.arange(rank)] = 1.0
    M[:rank, rank:] = rng.random((rank, n - rank))
    M[:, 24] = 0.0
    perm = np.arange(n, dtype=np.int32)
    k_js, k_0 = dbg.adapt_low_rank(M, np.zeros(n), np.ones(n), rank, perm, 1e-10)
    assert k_0 == 1 and perm[-1] == 24 and np.all(M[:, -1] == 0.0) and len(k_js) == rank


def test_leading_dimension_padding(dbg, oracle):
    """dbg_chol_rowmax_f64 and dbg_adapt_low_rank_f64 on a matrix stored with
    lda > n: the n x n block equals the unpadded results, the padding rows are
    untouched."""
    import ctypes

    from dbgenzpaper_amd import _lib
    from tests.test_oracle import _low_rank_case

    L = _lib.lib()
    n, pad = 90, 7
    C = random_spd(n, 77)
    a, b = -np.ones(n), np.ones(n)
    W = np.full((n + pad, n), 123.0, order="F")
    W[:n, :] = C
    av, bv = a.copy(), b.copy()
    perm = np.zeros(n, np.int32)
    scale = np.zeros(n)
    rank = ctypes.c_int32()
    dp = lambda x: _lib.ptr(x, ctypes.c_double)  # noqa: E731
    _lib.check(L.dbg_chol_rowmax_f64(dp(W), n + pad, n, dp(av), dp(bv), 16, 64, 128, 256,
                                     _lib.ptr(perm, ctypes.c_int32), dp(scale), ctypes.byref(rank)))
    F = dbg.cholesky_rowmax(C.copy(order="F"), a.copy(), b.copy(), 16, 64, 128, 256)
    assert rank.value == F.rank == n and np.array_equal(perm, F.perm)
    assert np.array_equal(W[:n, :], F.U) and np.all(W[n:, :] == 123.0)

    M, a2, b2 = _low_rank_case(n, 40, 8, (50, 89))
    Wm = np.full((n + pad, n), -5.0, order="F")
    Wm[:n, :] = M
    pv = np.arange(n, dtype=np.int32)
    kjs = np.zeros(40, np.int32)
    k0 = ctypes.c_int32()
    a3, b3 = a2.copy(), b2.copy()
    _lib.check(L.dbg_adapt_low_rank_f64(dp(Wm), n + pad, n, dp(a3), dp(b3), 40, _lib.ptr(pv, ctypes.c_int32), 1e-10,
                                        _lib.ptr(kjs, ctypes.c_int32), ctypes.byref(k0)))
    Mo, ao, bo, po, ko, k0o = oracle.adapt_low_rank(M, a2, b2, 40, np.arange(n), 1e-10)
    assert np.array_equal(Wm[:n, :], Mo) and np.all(Wm[n:, :] == -5.0)
    assert np.array_equal(pv, po) and np.array_equal(kjs, ko) and k0.value == k0o


def _copies_case(n, n_orig, seed, corr=True, big=1e6):
    """n_orig variables (correlated, one with variance `big`, so that
    eps0 = eps * max(diag) sits far above the rounding-level Schur diagonals
    of exact copies; or independent with exact unit correlations to their
    copies) followed by n - n_orig exact copies of them."""
    rng = np.random.default_rng(seed)
    if corr:
        A = rng.standard_normal((n_orig, n_orig))
        S0 = A @ A.T + n_orig * np.eye(n_orig)
    else:
        # variances powers of 4: the standardized copies correlate exactly 1
        S0 = np.diag(4.0 ** rng.integers(-1, 2, n_orig))
    S0[0, :] *= np.sqrt(big)
    S0[:, 0] *= np.sqrt(big)
    src = np.concatenate([np.arange(n_orig), rng.integers(0, n_orig, n - n_orig)])
    S = np.asfortranarray(S0[np.ix_(src, src)])
    sd = np.sqrt(np.diag(S))
    a = -rng.uniform(0.5, 2.0, n_orig)[src] * sd
    b = rng.uniform(0.5, 2.0, n_orig)[src] * sd
    return S, a, b


def test_genz_rank_stop_levels(dbg, oracle):
    """cholesky_genz! stopping for rank inside the second pivot window of the
    first bs3 block (n = 1500, bs2 = 512, bs3 = 1024): like the reference's
    break out of the loop nest (src/cholesky.jl:533-538), nothing later is
    solved -- the first window's rows stay unsolved past column 1024 (its
    solve reaches the end of its bs3 block only) -- and the whole upper
    triangle, the permutation and the rank equal the oracle's."""
    S, a, b = _copies_case(1500, 700, 7)
    print("genz rank stop n=1500 rank 700", flush=True)
    bs = (64, 512, 1024, 2048)
    Fo = oracle.cholesky_genz(S, a, b, *bs)
    Fg = dbg.cholesky_genz(S.copy(order="F"), a.copy(), b.copy(), *bs)
    print(f"  rank device {Fg.rank} oracle {Fo.rank}", flush=True)
    assert Fg.rank == Fo.rank == 700
    assert np.array_equal(Fg.perm, Fo.perm)
    du = np.abs(np.triu(Fg.U) - np.triu(Fo.U)).max()
    print(f"  max |dM| upper {du:.2e}", flush=True)
    assert du < 1e-12


def test_rowmax_rank_stop_windows(dbg, oracle):
    """cholesky_rowmax! with rank 300 < bs2 = 512 < n = 700 (default blocks):
    independent variables and exact copies, so the Schur diagonals of the
    copies are exactly 0 and the stop (m_ii <= 0 before the swap,
    src/cholesky.jl:329-350) falls at 300 in both; nothing after it is
    solved.  The whole upper triangle, permutation and rank equal the
    oracle's."""
    S, a, b = _copies_case(700, 300, 8, corr=False, big=1.0)
    print("rowmax rank stop n=700 rank 300", flush=True)
    Fo = oracle.cholesky_rowmax(S, a, b)
    Fg = dbg.cholesky_rowmax(S.copy(order="F"), a.copy(), b.copy())
    assert Fg.rank == Fo.rank == 300
    assert np.array_equal(Fg.perm, Fo.perm)
    assert np.abs(np.triu(Fg.U) - np.triu(Fo.U)).max() < 1e-12
