"""Synthetic workloads of the benchmark and the tests.

Each generator builds one problem family with its sizes, from a numpy
``default_rng(seed)`` stream (PCG64), so the same seed gives the same inputs
here and on any machine running this image.  None of them touches the GPU.

  dense_random(d)        Sigma = G G^T rescaled to unit diagonal (G d x d iid
                         N(0,1)), a = -inf, b ~ U(0, 2)
  factor_model(d, k)     Sigma = standardize(B B^T + diag(psi)), B d x k iid
                         N(0,1), psi ~ U(0.2, 1); a = -inf,
                         b = Phi^{-1}(u), u ~ U(lo, hi) (lower orthant)
  banded(d, bw)          L lower triangular with bandwidth bw, diag ~ U(1, 2),
                         in-band entries N(0, 0.3^2), rows scaled to unit norm,
                         Sigma = L L^T; a ~ U(-3, -1), b ~ U(1, 3)
  batch_uneven(K)        K independent dense_random problems, d_k ~ U{2..256},
                         b ~ U(-0.5, 2.5), in random arrival order
  equicorrelated(d)      Sigma = alpha I + beta 1 1^T (the reference's
                         fixed_dense, scripts/simulations/simulations_functions.jl:25-28)
                         with symmetric bounds of scripts/simulations/simulation_timed.jl:385-389
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.special import ndtri


@dataclass
class Problem:
    name: str
    Sigma: np.ndarray
    a: np.ndarray
    b: np.ndarray
    m: int
    n_reps: int
    shifts: np.ndarray | None = None
    meta: dict = field(default_factory=dict)

    @property
    def d(self) -> int:
        return self.Sigma.shape[0]


def _standardize(S: np.ndarray) -> np.ndarray:
    s = np.sqrt(np.diag(S))
    S = S / np.outer(s, s)
    np.fill_diagonal(S, 1.0)
    return np.asfortranarray((S + S.T) / 2)


def _shifts(rng, d, n_reps, gen):
    if gen == "Sobol":
        return np.asfortranarray(rng.integers(0, 2 ** 64, size=(d - 1, n_reps), dtype=np.uint64))
    return np.asfortranarray(rng.random((d - 1, n_reps)))


def dense_random(d=16, *, m=2 ** 14, n_reps=8, seed=0, gen="Richtmyer", b_range=(0.0, 2.0)) -> Problem:
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((d, d))
    Sigma = _standardize(G @ G.T)
    a = np.full(d, -np.inf)
    b = rng.uniform(b_range[0], b_range[1], d)
    return Problem(f"dense_random_d{d}", Sigma, a, b, m, n_reps, _shifts(rng, d, n_reps, gen),
                   {"gen": gen, "seed": seed})


def factor_model(d=500, k=8, *, m=2 ** 18, n_reps=10, seed=1, gen="Richtmyer", u_range=(0.01, 0.10)) -> Problem:
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((d, k))
    psi = rng.uniform(0.2, 1.0, d)
    Sigma = _standardize(B @ B.T + np.diag(psi))
    a = np.full(d, -np.inf)
    b = ndtri(rng.uniform(u_range[0], u_range[1], d))
    return Problem(f"factor_model_d{d}_k{k}", Sigma, a, b, m, n_reps, _shifts(rng, d, n_reps, gen),
                   {"gen": gen, "seed": seed, "k": k})


def banded(d=4000, bw=16, *, m=2 ** 18, n_reps=10, seed=2, gen="Richtmyer") -> Problem:
    rng = np.random.default_rng(seed)
    Lm = np.zeros((d, d))
    idx = np.arange(d)
    Lm[idx, idx] = rng.uniform(1.0, 2.0, d)
    for k in range(1, bw + 1):
        Lm[idx[k:], idx[:-k]] = rng.normal(0.0, 0.3, d - k)
    Lm /= np.linalg.norm(Lm, axis=1, keepdims=True)
    Sigma = Lm @ Lm.T
    np.fill_diagonal(Sigma, 1.0)
    Sigma = np.asfortranarray((Sigma + Sigma.T) / 2)
    a = rng.uniform(-3.0, -1.0, d)
    b = rng.uniform(1.0, 3.0, d)
    return Problem(f"banded_d{d}_bw{bw}", Sigma, a, b, m, n_reps, _shifts(rng, d, n_reps, gen),
                   {"gen": gen, "seed": seed, "bw": bw})


def batch_uneven(K=4096, *, d_min=2, d_max=256, m=2 ** 14, n_reps=4, seed=3):
    rng = np.random.default_rng(seed)
    dims = rng.integers(d_min, d_max + 1, size=K)
    probs = []
    for k, d in enumerate(dims):
        G = rng.standard_normal((d, d))
        Sigma = _standardize(G @ G.T)
        a = np.full(d, -np.inf)
        b = rng.uniform(-0.5, 2.5, d)
        probs.append(Problem(f"batch_{k}_d{d}", Sigma, a, b, m, n_reps, _shifts(rng, d, n_reps, "Richtmyer"),
                             {"index": k}))
    order = rng.permutation(K)
    return [probs[i] for i in order]


def equicorrelated(d, *, alpha=1.0, beta=1.0, m=2 ** 11, n_reps=12, seed=5, gen="Richtmyer") -> Problem:
    """fixed_dense (jitter alpha, common variance beta) with the symmetric
    bounds +-k*sqrt(diag), k = Phi^{-1}((1 + 0.25^(1/d)) / 2)."""
    rng = np.random.default_rng(seed)
    Sigma = np.asfortranarray(alpha * np.eye(d) + beta * np.ones((d, d)))
    k = ndtri((1 + 0.25 ** (1.0 / d)) / 2)
    s = np.sqrt(np.diag(Sigma))
    return Problem(f"equicorrelated_d{d}", Sigma, -k * s, k * s, m, n_reps,
                   _shifts(rng, d, n_reps, gen), {"alpha": alpha, "beta": beta, "gen": gen})


def equicorrelated_truth(alpha, beta, lower, upper) -> float:
    """P(lower <= X <= upper) for X ~ N(0, alpha I + beta 1 1^T) by 1-D
    quadrature over the common factor (the closed form of
    scripts/simulations/mwp.jl:80-116), in log space per point."""
    from scipy import integrate
    from scipy.special import log_ndtr

    lower = np.asarray(lower, float)
    upper = np.asarray(upper, float)
    isa = 1.0 / np.sqrt(alpha)
    sb = np.sqrt(beta)

    def logdiff(hi, lo):
        # log(Phi(hi) - Phi(lo)) without cancellation
        out = np.empty_like(hi)
        neg = hi <= 0
        lh, ll = log_ndtr(hi[neg]), log_ndtr(lo[neg])
        out[neg] = lh + np.log1p(-np.exp(ll - lh))
        pos = ~neg
        lh, ll = log_ndtr(-lo[pos]), log_ndtr(-hi[pos])
        out[pos] = lh + np.log1p(-np.exp(ll - lh))
        return out

    def f(z):
        c = sb * z
        return np.exp(np.sum(logdiff((upper - c) * isa, (lower - c) * isa))) * np.exp(-z * z / 2) / np.sqrt(2 * np.pi)

    val, _ = integrate.quad(f, -12, 12, epsabs=1e-15, epsrel=1e-13, limit=400)
    return float(val)
