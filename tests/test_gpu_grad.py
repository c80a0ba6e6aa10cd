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


@pytest.mark.parametrize("n,gen", [(64, "Richtmyer"), (200, "Richtmyer"), (300, "Sobol")])
def test_grad_mixed_close_to_oracle(dbg, oracle, n, gen):
    """Mixed forward and sweep (FP64 sums; from d = 96 the split-FP16 matrix-core
    kernel in record and reverse modes and the split Ubar product) against
    the FP64 oracle on the same factor, points and shifts: log P within 1e-6
    relative, the gradients within 1e-5 in norm (measured: log P 6e-9 to
    3e-8, gradients 9e-7 to 2.1e-6)."""
    print(f"grad mixed n={n} {gen}", flush=True)
    P = W.factor_model(n, 4, m=1024, n_reps=2, seed=n, u_range=(0.05, 0.3), gen=gen)
    D = _plan(dbg, P.Sigma, P.a, P.b, np.float32, 1024, 2, gen)
    D.set_shifts(P.shifts)
    g = D.grad()
    ro = _oracle_grad(oracle, D, gen)
    print(f"  log P rel {abs(g.log_value - ro.log_value) / abs(ro.log_value):.2e}, "
          f"b {_relnorm(g.b_std, ro.gb):.2e}, U {_relnorm(np.triu(g.U), np.triu(ro.gU)):.2e}", flush=True)
    assert abs(g.log_value - ro.log_value) < 1e-6 * abs(ro.log_value)
    assert _relnorm(g.b_std, ro.gb) < 1e-5
    assert _relnorm(np.triu(g.U), np.triu(ro.gU)) < 1e-5


@pytest.mark.parametrize("n,T,m,R", [(40, np.float64, 700, 3), (150, np.float64, 1000, 3),
                                     (64, np.float32, 1000, 3), (200, np.float32, 2000, 3)])
def test_grad_chunks_equal_one_chunk(dbg, oracle, n, T, m, R):
    """A record budget that splits the round into chunks of samples (the last
    one shorter, m not a multiple of the tile) gives the gradient of one
    chunk: FP64 to rounding, mixed within the FP32 rounding of the chunk-
    relative weights.  FP64 also against the oracle."""
    print(f"grad chunks n={n} {np.dtype(T).name} m={m} R={R}", flush=True)
    P = W.factor_model(n, 4, m=m, n_reps=R, seed=n + 1, u_range=(0.05, 0.3))
    D = _plan(dbg, P.Sigma, P.a, P.b, T, m, R)
    D.set_shifts(P.shifts)
    g1 = D.grad()
    rs1, nc1, bytes1 = D.grad_info()
    assert nc1 == 1 and rs1 >= m * R and bytes1 > 0
    rec_total = 2 * (n - 1) * rs1 * np.dtype(T).itemsize
    D.set_grad_budget(rec_total // 4)
    gc = D.grad()
    rs, nc, nbytes = D.grad_info()
    print(f"  chunks {nc} of {rs} samples, {nbytes / 2**20:.1f} MiB (one chunk: {bytes1 / 2**20:.1f} MiB)",
          flush=True)
    assert nc >= 3 and rs * nc >= m * R and nbytes < bytes1
    tol = 1e-11 if T == np.float64 else 1e-4
    assert abs(gc.log_value - g1.log_value) <= tol * abs(g1.log_value)
    assert _relnorm(gc.b_std, g1.b_std) < tol
    assert _relnorm(np.triu(gc.U), np.triu(g1.U)) < tol
    assert _relnorm(gc.Sigma, g1.Sigma) < tol
    if T == np.float64:
        ro = _oracle_grad(oracle, D, "Richtmyer")
        assert _relnorm(gc.b_std, ro.gb) < 1e-9
        assert _relnorm(np.triu(gc.U), np.triu(ro.gU)) < 1e-9
    D.set_grad_budget(0)
    assert D.grad_info() == (0, 0, 0)
    D.close()


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


def test_grad_c4_central_differences(dbg):
    """The gradient workload (factor model d = 1000, B 1000 x 16, lower
    orthant b = Phi^{-1}(u), u ~ U(0.02, 0.20); points cut to 2^13 x 8 for
    the test): the mixed and the FP64 gradients against FP64 central
    differences (h = 1e-5) of the device estimate on the same points, for 32
    random coordinates of b and 32 of Sigma (pivot order held: a coordinate
    whose perturbation changes it is skipped).  Bounds: 2e-6 (FP64) and
    1e-4 (mixed) of max(|fd|, 1); measured 2.7e-7 and 9.4e-6."""
    d, m, R, h = 1000, 2 ** 13, 8, 1e-5
    P = W.factor_model(d, 16, m=m, n_reps=R, seed=4, u_range=(0.02, 0.20))
    print("grad c4 fd", flush=True)

    def plan(S, b, T=np.float64):
        opts = dbg.QMC_opts(T, m=m, n_reps=R, max_pts=m, max_abs_err=0.0)
        return dbg.QMCData(S, P.a, b, opts, None, "Richtmyer", shifts=P.shifts, phi_form="tail")

    def logp(S, b):
        D = plan(S, b)
        dbg.qmc_pnorm(D)
        lp = D.last_result.log_value
        perm = D.C.perm.copy()
        D.close()
        return lp, perm

    D64 = plan(P.Sigma, P.b)
    perm0 = D64.C.perm.copy()
    g64 = D64.grad()
    g32 = plan(P.Sigma, P.b, np.float32).grad()
    rng = np.random.default_rng(7)
    checked = 0
    worst64, worst32 = 0.0, 0.0
    for i in rng.choice(d, 32, replace=False):
        e = np.zeros(d)
        e[i] = h
        lp_p, pp = logp(P.Sigma, P.b + e)
        lp_m, pm = logp(P.Sigma, P.b - e)
        if not (np.array_equal(pp, perm0) and np.array_equal(pm, perm0)):
            continue
        fd = (lp_p - lp_m) / (2 * h)
        scale = max(abs(fd), 1.0)
        worst64 = max(worst64, abs(g64.b[i] - fd) / scale)
        worst32 = max(worst32, abs(g32.b[i] - fd) / scale)
        checked += 1
    for _ in range(32):
        i, j = rng.choice(d, 2, replace=False)
        E = np.zeros((d, d))
        E[i, j] = E[j, i] = h
        lp_p, pp = logp(P.Sigma + E, P.b)
        lp_m, pm = logp(P.Sigma - E, P.b)
        if not (np.array_equal(pp, perm0) and np.array_equal(pm, perm0)):
            continue
        fd = (lp_p - lp_m) / (2 * h)
        scale = max(abs(fd), 1.0)
        worst64 = max(worst64, abs(2 * g64.Sigma[i, j] - fd) / scale)
        worst32 = max(worst32, abs(2 * g32.Sigma[i, j] - fd) / scale)
        checked += 1
    print(f"c4 fd: {checked} coordinates, worst fp64 {worst64:.2e}, mixed {worst32:.2e}", flush=True)
    assert checked >= 48
    assert worst64 < 2e-6
    assert worst32 < 1e-4
