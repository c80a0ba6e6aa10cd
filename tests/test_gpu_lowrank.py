"""Dense plans on a rank-deficient Sigma (no reference counterpart: QMCData
throws for it, src/qmc.jl:195-197): the device route (row-max factor with a
rank tolerance, dependent columns completed by dtrsm, adapt_low_rank!, FP64
level loop) against the oracle's route on the same Sigma, a, b and shifts;
rank 1 against its closed form."""
import numpy as np
import pytest
from scipy.stats import norm

from dbgenzpaper_amd import workloads as W
from dbgenzpaper_amd.sobol import direction_integers, skip_count
from tests._helpers import rel

pytestmark = pytest.mark.gpu


def _mixed_case(seed=21):
    """Twelve variables on eight latent ones: a copy, a negated multiple, a
    combination of two and a generic combination."""
    rng = np.random.default_rng(seed)
    r = 8
    B = rng.standard_normal((r, r))
    Sr = B @ B.T + r * np.eye(r)
    T = np.zeros((12, r))
    T[:r] = np.eye(r)
    T[8, 2] = 1.0
    T[9, 5] = -2.5
    T[10, [1, 3]] = [0.7, -1.1]
    T[11] = rng.standard_normal(r)
    S = T @ Sr @ T.T
    sd = np.sqrt(np.diag(S))
    a, b = -rng.uniform(0.4, 2.0, 12) * sd, rng.uniform(0.4, 2.0, 12) * sd
    a[[0, 6]] = -np.inf
    return S, a, b


def _device_vs_oracle(dbg, oracle, S, a, b, T, gen, phi, m, R, seed, utol=1e-11):
    n = S.shape[0]
    rng = np.random.default_rng(seed)
    LF = oracle.prepare_lowrank(S, a, b)
    if gen == "Sobol":
        sh = np.asfortranarray(rng.integers(0, 2 ** 64, size=(n - 1, R), dtype=np.uint64))
        kw = dict(V=direction_integers(n - 1), sobol_skip=skip_count(2 * m - 1))
    else:
        sh = np.asfortranarray(rng.random((n - 1, R)))
        kw = {}
    ro = oracle.qmc_pnorm_lowrank(LF, m=m, n_reps=R, gen=gen, phi_form=phi, shifts=sh, **kw)
    opts = dbg.QMC_opts(T, m=m, n_reps=R, max_pts=m, max_abs_err=0.0)
    D = dbg.QMCData(S, a, b, opts, None, gen, shifts=sh, phi_form=phi)
    r, kjs, k0 = D.lowrank_info()
    assert r == LF.rank and k0 == LF.k0 and np.array_equal(kjs + 1, LF.beg[1:])
    F = D.C
    assert F.rank == LF.rank and np.array_equal(F.perm, LF.perm)
    # per column, relative to its largest entry: adapt_low_rank! divides a
    # column by its last coefficient, which scales the completion's rounding
    colmax = np.maximum(np.abs(LF.U[:r]).max(axis=0), 1.0)
    assert (np.abs(F.U[:r] - LF.U[:r]) / colmax).max() < utol
    v, e, _ = dbg.qmc_pnorm(D)
    reps = np.array(D.qmc_reps)
    D.close()
    return v, ro, reps


@pytest.mark.parametrize("gen,phi", [("Richtmyer", "erf"), ("Richtmyer", "tail"), ("Sobol", "erf"),
                                     ("Richtmyer", "erfc")])
def test_mixed_dependencies_match_oracle(dbg, oracle, gen, phi):
    S, a, b = _mixed_case()
    print(f"lowrank n=12 r=8 {gen} {phi}", flush=True)
    v, ro, reps = _device_vs_oracle(dbg, oracle, S, a, b, np.float64, gen, phi, 1500, 4, 22)
    print(f"  device {v:.15e} oracle {ro.value:.15e}", flush=True)
    assert rel(v, ro.value) < 1e-11
    assert np.allclose(reps, ro.rep_means, rtol=1e-11, atol=0)


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_exact_factor_model_matches_oracle(dbg, oracle, T):
    """Sigma = B B' with B 300 x 40 (no diagonal term): rank 40, every variable
    a generic combination, the dependent ones all at the last level.  A mixed
    plan takes the same FP64 route.  The last coefficients of some dependent
    variables are of order 1e-4, so the scaled columns reach 4e3 and carry
    the completion's rounding (rocBLAS and LAPACK solves) amplified by 1e4:
    factor and estimate to 1e-9."""
    rng = np.random.default_rng(23)
    n, k = 300, 40
    B = rng.standard_normal((n, k))
    S = B @ B.T
    sd = np.sqrt(np.diag(S))
    a, b = -np.inf * np.ones(n), rng.uniform(2.0, 4.0, n) * sd
    print(f"lowrank factor model n={n} k={k} {np.dtype(T).name}", flush=True)
    v, ro, reps = _device_vs_oracle(dbg, oracle, S, a, b, T, "Richtmyer", "tail", 2048, 4, 24, utol=1e-9)
    print(f"  device {v:.15e} oracle {ro.value:.15e}", flush=True)
    assert ro.value > 0.1 and rel(v, ro.value) < 1e-9
    assert np.allclose(reps, ro.rep_means, rtol=1e-9, atol=0)


def test_rank_one_closed_form(dbg):
    """X_j = s_j Z: P = Phi(min hi_j) - Phi(max lo_j) with nothing sampled."""
    s = np.array([1.0, 2.0, -1.5, 0.5, -3.0, 1.0, 0.25])
    S = np.outer(s, s)
    a = np.array([-1.0, -3.0, -2.0, -0.4, -4.0, -np.inf, -1.0])
    b = np.array([1.5, 2.5, 1.0, 1.0, 5.0, 0.8, 0.6])
    lo, hi = np.where(s > 0, a / s, b / s), np.where(s > 0, b / s, a / s)
    exact = norm.cdf(hi.min()) - norm.cdf(lo.max())
    print("lowrank rank 1", flush=True)
    opts = dbg.QMC_opts(np.float64, m=256, n_reps=3, max_pts=256, max_abs_err=0.0)
    D = dbg.QMCData(S, a, b, opts, None, "Richtmyer", shifts=np.zeros((6, 3)))
    assert D.lowrank_info()[0] == 1
    v, e, _ = dbg.qmc_pnorm(D)
    D.close()
    assert rel(v, exact) < 1e-14 and e < 1e-14 * v


def test_refactor_and_gradient_guard(dbg):
    """dbg_plan_refactor takes the same route again; the gradient is refused
    for a rank-deficient plan."""
    S, a, b = _mixed_case(25)
    opts = dbg.QMC_opts(np.float64, m=1024, n_reps=2, max_pts=1024, max_abs_err=0.0)
    D = dbg.QMCData(S, a, b, opts, np.random.default_rng(26), "Richtmyer")
    v1, _, _ = dbg.qmc_pnorm(D)
    D.gen_state = 0
    D.refactor()
    v2, _, _ = dbg.qmc_pnorm(D)
    assert v1 == v2 and D.lowrank_info()[0] == 8
    with pytest.raises(Exception):
        D.grad()
    D.close()


def test_full_rank_plans_unchanged(dbg):
    """A Sigma of full rank keeps the Genz route: rank n, one column per level."""
    P = W.dense_random(20, m=512, n_reps=2, seed=27)
    opts = dbg.QMC_opts(np.float64, m=512, n_reps=2, max_pts=512)
    D = dbg.QMCData(P.Sigma, P.a, P.b, opts, None, "Richtmyer", shifts=P.shifts)
    r, kjs, k0 = D.lowrank_info()
    assert r == 20 and np.array_equal(kjs, np.arange(20)) and k0 == 0 and D.C.rank == 20
    D.close()
