"""The oracle's route for a rank-deficient Sigma (no reference counterpart:
QMCData throws for it, src/qmc.jl:195-197) pinned by what the mathematics
fixes: with one column per block the level loop is the dense loop bit for
bit; rank 1 is a closed form; rank 2 a one-dimensional integral; exact copies
of variables reduce to the full-rank problem with intersected bounds; a
generic rank-deficient problem agrees with direct simulation; the completed
factor reproduces the correlations of every pair of variables."""
import numpy as np
import pytest
from scipy import integrate
from scipy.stats import norm

from dbgenzpaper_amd.sobol import direction_integers, skip_count


def _shifts(rng, n, R, gen):
    if gen == "Sobol":
        return np.asfortranarray(rng.integers(0, 2 ** 32, size=(n - 1, R), dtype=np.uint64))
    return np.asfortranarray(rng.random((n - 1, R)))


@pytest.mark.parametrize("phi_form", ["erf", "erfc", "tail"])
@pytest.mark.parametrize("gen", ["Richtmyer", "Sobol"])
def test_level_loop_is_the_dense_loop_at_full_rank(oracle, phi_form, gen):
    """One column per block (beg = 0 .. n): the level loop draws and sums as
    orc_qmc_loop_dense does (same fma order of the conditional means, same
    batches), so the shift means agree bit for bit."""
    rng = np.random.default_rng(5)
    n, m, R = 30, 1000, 4
    B = rng.standard_normal((n, n))
    S = B @ B.T + n * np.eye(n)
    a, b = -rng.uniform(0.5, 2, n), rng.uniform(0.5, 2, n)
    a[::3] = -np.inf
    F = oracle.prepare_dense(S, a, b)
    sh = _shifts(rng, n, R, gen)
    kw = dict(V=direction_integers(n - 1), sobol_skip=skip_count(2 * m - 1)) if gen == "Sobol" else {}
    rd = oracle.qmc_pnorm(F.U, F.a, F.b, m=m, n_reps=R, gen=gen, shifts=sh, phi_form=phi_form, bs_i=8, bs_i2=4,
                          **kw)
    LF = oracle.LowRankFactor(F.U, F.a, F.b, F.perm, n, np.arange(n + 1, dtype=np.int32), 0,
                              1.0 / (np.diag(F.U) * np.sqrt(2.0)), F.scale_vec)
    rl = oracle.qmc_pnorm_lowrank(LF, m=m, n_reps=R, gen=gen, shifts=sh, phi_form=phi_form, **kw)
    assert np.array_equal(rd.rep_mant, rl.rep_mant) and np.array_equal(rd.rep_exp, rl.rep_exp)
    assert rd.value == rl.value


def test_rank_one_closed_form(oracle):
    """X_j = s_j Z: P = Phi(min hi_j) - Phi(max lo_j), [lo_j, hi_j] the
    interval of Z from constraint j (exchanged for s_j < 0); nothing is
    sampled at rank 1."""
    s = np.array([1.0, 2.0, -1.5, 0.5, -3.0, 1.0])
    S = np.outer(s, s)
    a = np.array([-1.0, -3.0, -2.0, -0.4, -4.0, -np.inf])
    b = np.array([1.5, 2.5, 1.0, 1.0, 5.0, 0.8])
    lo, hi = np.where(s > 0, a / s, b / s), np.where(s > 0, b / s, a / s)
    exact = norm.cdf(hi.min()) - norm.cdf(lo.max())
    LF = oracle.prepare_lowrank(S, a, b)
    assert LF.rank == 1 and list(LF.beg) == [0, 6] and LF.k0 == 0
    r = oracle.qmc_pnorm_lowrank(LF, m=64, n_reps=4, shifts=np.zeros((5, 4)))
    assert abs(r.value / exact - 1) < 1e-14 and r.err3 == 0.0


def _rank_two_case():
    rho, al, be = 0.4, 0.7, -1.2
    s2 = np.sqrt(1 - rho ** 2)
    L = np.array([[1, 0], [rho, s2], [al + be * rho, be * s2]])
    a, b = np.array([-1.0, -0.5, -1.5]), np.array([1.2, 2.0, 0.9])

    def integrand(z1):
        lo2, hi2 = (a[1] - rho * z1) / s2, (b[1] - rho * z1) / s2
        c3, m3 = be * s2, (al + be * rho) * z1
        l3, h3 = sorted(((a[2] - m3) / c3, (b[2] - m3) / c3))
        return norm.pdf(z1) * max(norm.cdf(min(hi2, h3)) - norm.cdf(max(lo2, l3)), 0.0)

    exact, _ = integrate.quad(integrand, a[0], b[0], epsabs=1e-14, epsrel=1e-12, limit=200)
    return L @ L.T, a, b, exact


def test_rank_two_quadrature(oracle):
    """X3 = 0.7 X1 - 1.2 X2: latent 2's interval is the intersection of
    the constraints of X2 and X3 (exchanged, the coefficient being negative);
    against the one-dimensional integral by adaptive quadrature, within the
    estimate's own error."""
    S, a, b, exact = _rank_two_case()
    LF = oracle.prepare_lowrank(S, a, b)
    assert LF.rank == 2 and list(LF.beg) == [0, 1, 3]
    rng = np.random.default_rng(6)
    r = oracle.qmc_pnorm_lowrank(LF, m=2 ** 16, n_reps=16, shifts=np.asfortranarray(rng.random((2, 16))))
    assert abs(r.value - exact) < max(r.err3, 1e-9) and r.err3 < 1e-6 * exact


def test_zero_variance_columns(oracle):
    """Columns with no coefficient at the end (beg[r] < n) multiply the
    estimate by [a_c <= 0 <= b_c]."""
    S, a, b, _ = _rank_two_case()
    LF = oracle.prepare_lowrank(S, a, b)
    sh = np.asfortranarray(np.random.default_rng(7).random((3, 4)))
    base = oracle.qmc_pnorm_lowrank(LF, m=4096, n_reps=4, shifts=sh[:2]).value
    U = np.zeros((4, 4), order="F")
    U[:3, :3] = LF.U
    for (a3, b3), want in (((-1.0, 1.0), base), ((0.2, 1.0), 0.0), ((-np.inf, -0.1), 0.0)):
        LZ = oracle.LowRankFactor(U, np.r_[LF.a, a3], np.r_[LF.b, b3], np.r_[LF.perm, 3], 2, LF.beg, 1,
                                  np.r_[LF.inv, 0.0], np.r_[LF.scale_vec, 1.0])
        assert oracle.qmc_pnorm_lowrank(LZ, m=4096, n_reps=4, shifts=sh).value == want


def test_copies_reduce_to_the_full_rank_problem(oracle):
    """Variables that are multiples c X_k of others (c > 0 and c < 0): each
    joins the block of its latent variable with coefficient row U[:, k] / U_kk
    and bounds [a, b] / (c U_kk), and the estimate is the dense loop's on the
    latent factor with the intersected bounds, on the same points."""
    rng = np.random.default_rng(8)
    r = 6
    B = rng.standard_normal((r, r))
    Sr = B @ B.T + r * np.eye(r)
    src, mult = [1, 4, 1, 0, 5], [2.0, -0.5, -1.0, 1.5, 3.0]
    T = np.zeros((r + len(src), r))
    T[:r] = np.eye(r)
    for q, (k, c) in enumerate(zip(src, mult)):
        T[r + q, k] = c
    S = T @ Sr @ T.T
    n = S.shape[0]
    sd = np.sqrt(np.diag(S))
    a, b = -rng.uniform(0.3, 1.5, n) * sd, rng.uniform(0.3, 1.5, n) * sd
    LF = oracle.prepare_lowrank(S, a, b)
    assert LF.rank == r and LF.k0 == 0
    prim = LF.beg[:r]
    Ur = np.asfortranarray(LF.U[:r][:, prim])
    ar, br = LF.a[prim].copy(), LF.b[prim].copy()
    for i in range(r):
        uii = Ur[i, i]
        for c in range(LF.beg[i] + 1, LF.beg[i + 1]):
            assert np.allclose(LF.U[:i + 1, c], Ur[:i + 1, i] / uii, rtol=0, atol=1e-12)
            ar[i], br[i] = max(ar[i], LF.a[c] * uii), min(br[i], LF.b[c] * uii)
    sh = np.asfortranarray(rng.random((n - 1, 8)))
    rl = oracle.qmc_pnorm_lowrank(LF, m=2048, n_reps=8, shifts=sh)
    rd = oracle.qmc_pnorm(Ur, ar, br, m=2048, n_reps=8, shifts=np.asfortranarray(sh[:r - 1]), bs_i=r - 1,
                          bs_i2=r - 1)
    assert np.allclose(rl.rep_means, rd.rep_means, rtol=1e-12, atol=0)


def test_generic_rank_deficient_against_simulation(oracle):
    """Eight variables on four latent ones, every dependent variable a
    generic combination (all in the last block): the estimate against
    8e6 direct draws of X = L Z, within 4 standard errors of the two."""
    rng = np.random.default_rng(4)
    n, r = 8, 4
    L = rng.standard_normal((n, r))
    S = L @ L.T
    sd = np.sqrt(np.diag(S))
    a, b = -rng.uniform(0.3, 1.5, n) * sd, rng.uniform(0.3, 2.0, n) * sd
    LF = oracle.prepare_lowrank(S, a, b)
    assert LF.rank == r and list(LF.beg) == [0, 1, 2, 3, 8]
    q = oracle.qmc_pnorm_lowrank(LF, m=2 ** 14, n_reps=16, shifts=np.asfortranarray(rng.random((n - 1, 16))))
    hits, N = 0, 0
    for _ in range(4):
        X = rng.standard_normal((2_000_000, r)) @ L.T
        hits += int(np.all((X >= a) & (X <= b), axis=1).sum())
        N += 2_000_000
    p = hits / N
    se = np.hypot(np.sqrt(p * (1 - p) / N), q.err3 / 3)
    assert abs(q.value - p) < 4 * se


def test_completed_factor_reproduces_the_correlations(oracle):
    """The columns of the completed, reordered factor are, up to the scaling
    of adapt_low_rank!, the latent coefficients of the variables they hold:
    the cosine of any two columns is the correlation of their variables."""
    rng = np.random.default_rng(9)
    n, r = 40, 12
    L = rng.standard_normal((n, r)) * rng.uniform(0.5, 3.0, (n, 1))
    S = L @ L.T
    a, b = -np.ones(n), np.ones(n)
    LF = oracle.prepare_lowrank(S, a, b, chol_block_size=4, chol_block_size2=16)
    assert LF.rank == r
    V = LF.U[:r]
    cosv = (V.T @ V) / np.outer(np.linalg.norm(V, axis=0), np.linalg.norm(V, axis=0))
    sd = np.sqrt(np.diag(S))
    Cp = (S / np.outer(sd, sd))[np.ix_(LF.perm, LF.perm)]
    assert np.abs(np.abs(cosv) - np.abs(Cp)).max() < 1e-10
