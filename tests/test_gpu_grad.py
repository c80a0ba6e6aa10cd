"""Device gradient of log P (the adjoint; no reference counterpart) against
the oracle's FP64 reverse sweep on the same factor, points and shifts, and
the Sigma chain rule against the central-difference-pinned one of the tests.
"""
import numpy as np
import pytest

from dbgenzpaper_amd import workloads as W
from dbgenzpaper_amd.sobol import direction_integers, skip_count
from tests._helpers import chain_to_sigma, random_spd

pytestmark = pytest.mark.gpu


def _relnorm(x, y):
    x, y = np.asarray(x), np.asarray(y)
    return float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))


def _problem(n, seed, lower_inf=True):
    rng = np.random.default_rng(seed)
    S = random_spd(n, seed, jitter=float(n))
    a, b = -rng.uniform(0.3, 2.0, n), rng.uniform(0.3, 2.0, n)
    if lower_inf:
        a[rng.choice(n, size=max(1, n // 3), replace=False)] = -np.inf
    return S, a, b


def _oracle_grad(oracle, D, gen):
    F = D.C
    kw = {}
    if gen == "Sobol":
        kw = dict(V=direction_integers(D.n - 1), sobol_skip=skip_count(2 * D.qmc_opts.m - 1))
    return oracle.qmc_grad(F.U, F.a, F.b, m=D.qmc_opts.m, n_reps=D.qmc_opts.n_reps, gen=gen,
                           shifts=D.shifts, **kw)


def _plan(dbg, S, a, b, T, m, R, gen="Richtmyer", seed=0):
    opts = dbg.QMC_opts(T, m=m, n_reps=R, max_pts=m, max_abs_err=0.0)
    rng = np.random.default_rng(seed)
    return dbg.QMCData(S, a, b, opts, rng, gen, phi_form="tail")


@pytest.mark.parametrize("n,m,R,gen", [(12, 1000, 4, "Richtmyer"), (40, 700, 3, "Richtmyer"),
                                       (150, 512, 2, "Richtmyer"), (24, 512, 2, "Sobol")])
def test_grad_fp64_matches_oracle(dbg, oracle, n, m, R, gen):
    print(f"grad fp64 n={n} m={m} R={R} {gen}", flush=True)
    S, a, b = _problem(n, n)
    D = _plan(dbg, S, a, b, np.float64, m, R, gen)
    g = D.grad()
    ro = _oracle_grad(oracle, D, gen)
    assert abs(g.log_value - ro.log_value) < 1e-11 * abs(ro.log_value)
    assert _relnorm(g.b_std, ro.gb) < 1e-9
    assert _relnorm(g.a_std, ro.ga) < 1e-9 or np.all(np.isinf(D.C.a))
    assert _relnorm(np.triu(g.U), np.triu(ro.gU)) < 1e-9
    # Sigma, a, b through the fixed-order factorization
    ga, gb, G = chain_to_sigma(D.C.U, D.C.perm, S, ro.ga, ro.gb, ro.gU, D.C.a, D.C.b)
    assert _relnorm(g.Sigma, G) < 1e-8
    assert _relnorm(g.b, gb) < 1e-9 and _relnorm(g.a, ga) < 1e-9 or np.all(np.isinf(a))
    assert np.allclose(g.Sigma, g.Sigma.T, rtol=0, atol=1e-12 * np.abs(g.Sigma).max())


@pytest.mark.parametrize("n", [64, 200])
def test_grad_mixed_close_to_oracle(dbg, oracle, n):
    """FP32 forward and sweep (FP64 sums): within 2e-3 of the FP64 oracle in norm."""
    print(f"grad mixed n={n}", flush=True)
    P = W.factor_model(n, 4, m=1024, n_reps=2, seed=n, u_range=(0.05, 0.3))
    D = _plan(dbg, P.Sigma, P.a, P.b, np.float32, 1024, 2)
    g = D.grad()
    ro = _oracle_grad(oracle, D, "Richtmyer")
    assert abs(g.log_value - ro.log_value) < 1e-4 * abs(ro.log_value)
    assert _relnorm(g.b_std, ro.gb) < 2e-3
    assert _relnorm(np.triu(g.U), np.triu(ro.gU)) < 2e-3


def test_grad_repeatable_and_state(dbg):
    """Two calls on one plan give the same gradient; the forward estimate is
    unchanged by them."""
    S, a, b = _problem(30, 5)
    D = _plan(dbg, S, a, b, np.float64, 512, 2)
    v0, _, _ = dbg.qmc_pnorm(D)
    g1 = D.grad()
    g2 = D.grad()
    assert np.array_equal(g1.b_std, g2.b_std) or _relnorm(g1.b_std, g2.b_std) < 1e-14
    assert _relnorm(g1.Sigma, g2.Sigma) < 1e-13
    v1, _, _ = dbg.qmc_pnorm(D)
    assert v0 == v1
