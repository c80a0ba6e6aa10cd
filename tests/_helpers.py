"""Shared test helpers: run the CPU oracle on exactly the inputs a device
plan holds (same factor, bounds, points, shifts and generator state)."""
from __future__ import annotations

import numpy as np

from dbgenzpaper_amd import _lib as L
from dbgenzpaper_amd.sobol import direction_integers, skip_count


def oracle_same_inputs(O, D, *, kind=None, phi=None, gen_state=None, max_pts=None, max_abs_err=None,
                       U=None, a=None, b=None):
    """qmc_pnorm! of the oracle on the plan's own factor and shifts."""
    F = D.C
    opts = D.qmc_opts
    n = D.n
    if kind is None:
        kind = {L.KIND_DENSE: "dense", L.KIND_SPARSE: "sparse", L.KIND_SPARSEQMC: "sparseqmc"}[D._kind]
    if phi is None:
        phi = {L.PHI_ERF: "erf", L.PHI_ERFC: "erfc", L.PHI_TAIL: "tail"}[D._phi]
        if D.T == np.float32:
            phi = "tail"
    gen = "Sobol" if (D.qmc_type == "Sobol" or kind == "sparseqmc") else "Richtmyer"
    V = direction_integers(max(n - 1, 1)) if gen == "Sobol" else None
    skip = D._sobol_skip if D._sobol_skip is not None else skip_count(2 * opts.m - 1)
    bs_i, bs_i2 = max(opts.block_size_i, 1), max(opts.block_size_i2, 1)
    return O.qmc_pnorm(F.U if U is None else U, F.a if a is None else a, F.b if b is None else b,
                       m=opts.m, n_reps=opts.n_reps, gen=gen, phi_form=phi, kind=kind,
                       bs_i=bs_i, bs_i2=bs_i2, bs_j=opts.block_size_j, n_bits=opts.n_bits,
                       max_pts=opts.max_pts if max_pts is None else max_pts,
                       max_abs_err=opts.max_abs_err if max_abs_err is None else max_abs_err,
                       shifts=D.shifts, V=V, sobol_skip=skip, gen_state=gen_state)


def rel(x, y):
    return abs(x - y) / max(abs(y), 1e-300)


def random_spd(n, seed, jitter=100.0):
    """C = A A^T + jitter I with A ~ U(0,1)^{n x n} (test/runtests.jl:16-19)."""
    rng = np.random.default_rng(seed)
    A = rng.random((n, n))
    return np.asfortranarray(A @ A.T + jitter * np.eye(n))


def standardized(C):
    d = np.sqrt(np.diag(C))
    return C / np.outer(d, d)


def chain_to_sigma(U, perm, Sigma, ga_std, gb_std, gU, a_std, b_std):
    """Gradient of log P with respect to the original a, b and Sigma from the
    gradient with respect to the standardized, permuted bounds and the factor
    U (C_p = U^T U, permutation held fixed): the Cholesky adjoint
    C_bar = U^{-1} Phi(U U_bar^T) U^{-T} (Phi: lower triangle, halved
    diagonal; then symmetrized), the standardization C = D^{-1/2} Sigma D^{-1/2}
    and a' = a / sqrt(Sigma_ii).  Returns (ga, gb, G) with d log P =
    sum_ij G_ij E_ij for a symmetric perturbation E of Sigma."""
    n = U.shape[0]
    perm = np.asarray(perm)
    s = np.sqrt(np.diag(Sigma))
    sp = s[perm]
    Ub = np.triu(gU)
    X = U @ Ub.T
    P = np.tril(X)
    P[np.diag_indices(n)] *= 0.5
    Z = np.linalg.solve(U, P)
    Cb = np.linalg.solve(U, Z.T).T
    Cb = 0.5 * (Cb + Cb.T)
    Cp = U.T @ U
    G = np.zeros((n, n))
    G[np.ix_(perm, perm)] = Cb / np.outer(sp, sp)
    for i in range(n):
        k = perm[i]
        G[k, k] -= np.sum(Cb[i, :] * Cp[i, :]) / Sigma[k, k]
        ta = 0.0 if np.isinf(a_std[i]) else ga_std[i] * a_std[i]
        tb = 0.0 if np.isinf(b_std[i]) else gb_std[i] * b_std[i]
        G[k, k] -= (ta + tb) / (2 * Sigma[k, k])
    ga = np.zeros(n)
    gb = np.zeros(n)
    ga[perm] = ga_std / sp
    gb[perm] = gb_std / sp
    return ga, gb, G


def logp_fixed_perm(O, Sigma, a, b, perm, *, m, n_reps, shifts):
    """log P (tail form) with the factor of the permuted, standardized Sigma
    taken by a plain Cholesky in a fixed order: the function whose
    derivative the adjoint returns."""
    perm = np.asarray(perm)
    s = np.sqrt(np.diag(Sigma))
    Cp = (Sigma / np.outer(s, s))[np.ix_(perm, perm)]
    U = np.linalg.cholesky(Cp).T
    ap, bp = a[perm] / s[perm], b[perm] / s[perm]
    n = Sigma.shape[0]
    r = O.qmc_pnorm(U, ap, bp, m=m, n_reps=n_reps, shifts=shifts, phi_form="tail", bs_i=n - 1, bs_i2=n - 1)
    return r.log_value, U, ap, bp


def normal_quantile_mp(p, mpmath):
    """Phi^{-1}(p) for 0 < p <= 1/2 (an mpmath number) by Newton steps on
    log Phi(y) = log p, to about 30 digits."""
    y = mpmath.mpf(-0.5) if p > 0.1 else -mpmath.sqrt(-2 * mpmath.log(p))
    for _ in range(200):
        f = mpmath.log(mpmath.ncdf(y)) - mpmath.log(p)
        yn = y - f * mpmath.ncdf(y) / mpmath.npdf(y)
        if abs(yn - y) < mpmath.mpf(10) ** -30:
            return yn
        y = yn
    return y
