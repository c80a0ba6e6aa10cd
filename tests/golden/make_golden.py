#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden/ (run from the repository root:
``python tests/golden/make_golden.py``).

The reference ships no numeric vectors for this path, so the fixtures come
from three sources, recorded per file:

* ``truths.json`` -- values fixed by mathematics or by the reference's own
  data files: closed forms (Sigma = I), library routines (SciPy's bivariate
  normal CDF, mpmath quantiles at 40 digits), the 1-D quadrature truth of the
  equicorrelated family, and the fixed_dense medians the reference publishes
  in data/sims/sparse_dense_vals.csv;
* ``sobol_d8.npz`` -- the first 1024 unscrambled Sobol points of 8
  dimensions from SciPy (Joe-Kuo direction numbers), as 32-bit words;
* ``config0_d16.npz`` and ``genz_perm.npz`` -- inputs and the CPU oracle's
  outputs (factor, permutation, estimate) for the d = 16 headline CPU case and
  for small pivoted factorizations: regression vectors that the oracle must
  keep reproducing and that the device must match (permutation bit-exact).
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def truths() -> dict:
    import mpmath as mp
    from scipy.special import ndtr
    from scipy.stats import multivariate_normal

    from dbgenzpaper_amd import workloads as W

    mp.mp.dps = 40
    out = {}
    # Sigma = I: every sample's weight is prod(Phi(b) - Phi(a)) exactly
    rng = np.random.default_rng(3)
    a, b = -rng.uniform(0.2, 2, 16), rng.uniform(0.2, 2, 16)
    out["identity_d16"] = {"a": a.tolist(), "b": b.tolist(),
                           "value": float(np.prod(ndtr(b) - ndtr(a)))}
    # d = 2: the bivariate normal CDF
    S = [[1.0, 0.6], [0.6, 1.0]]
    bb = [0.3, -0.4]
    out["bivariate"] = {"Sigma": S, "b": bb,
                        "value": float(multivariate_normal([0, 0], S).cdf(bb))}
    # equicorrelated Sigma = 1 1^T + I (the reference's fixed_dense family)
    eq = {}
    for d in (16, 64, 256):
        P = W.equicorrelated(d, m=16)
        eq[str(d)] = {"a": P.a.tolist(), "b": P.b.tolist(),
                      "value": float(W.equicorrelated_truth(1.0, 1.0, P.a, P.b))}
    out["equicorrelated"] = eq
    # published fixed_dense medians, data/sims/sparse_dense_vals.csv (m = 245760)
    out["published_fixed_dense_medians"] = {"256": 0.698106405416119, "512": 0.7388865012733516}
    # normal quantiles Phi^{-1}(t 2^-k) at 40 digits
    qs = []
    for lp in np.concatenate([np.linspace(-1.0001, -40.0, 60), np.linspace(-40.0, -3000.0, 40)]):
        k = max(0, int(-lp) - 100) if lp < -110 else 0
        t = float(np.float32(2.0 ** (lp + k)))
        p = mp.mpf(t) * mp.mpf(2) ** (-k)
        y = mp.mpf(-0.5) if p > 0.1 else -mp.sqrt(-2 * mp.log(p))
        for _ in range(200):
            yn = y - (mp.log(mp.ncdf(y)) - mp.log(p)) * mp.ncdf(y) / mp.npdf(y)
            if abs(yn - y) < mp.mpf(10) ** -30:
                y = yn
                break
            y = yn
        qs.append([t, k, float(y)])
    out["quantiles_t_k_y"] = qs
    return out


def sobol_words() -> np.ndarray:
    from scipy.stats import qmc

    x = qmc.Sobol(d=8, scramble=False).random(1024)
    return np.floor(np.minimum(x, np.nextafter(1.0, 0.0)) * 2.0 ** 32).astype(np.uint32)


def config0(O) -> dict:
    from dbgenzpaper_amd import workloads as W

    P = W.dense_random(16)
    F = O.prepare_dense(P.Sigma, P.a, P.b)
    bs_i, bs_i2 = O.clamp_block_sizes(16, 2 ** 10, 2 ** 6)
    r = O.qmc_pnorm(F.U, F.a, F.b, m=P.m, n_reps=P.n_reps, shifts=P.shifts, bs_i=bs_i, bs_i2=bs_i2)
    return dict(Sigma=P.Sigma, a=P.a, b=P.b, shifts=P.shifts, m=P.m, n_reps=P.n_reps,
                U=F.U, perm=F.perm, a_perm=F.a, b_perm=F.b, rank=F.rank,
                value=r.value, err3=r.err3, log_value=r.log_value, rep_means=np.asarray(r.rep_means))


def genz_perms(O) -> dict:
    from tests._helpers import random_spd

    out = {}
    for i, (n, bs, seed) in enumerate([(12, (2, 4, 8, 8), 1), (37, (4, 8, 16, 32), 2),
                                       (64, (8, 32, 64, 64), 3), (100, (16, 64, 128, 128), 4)]):
        C = random_spd(n, seed)
        rng = np.random.default_rng(seed)
        a, b = -rng.uniform(0.5, 3, n), rng.uniform(0.5, 3, n)
        F = O.cholesky_genz(C, a, b, *bs)
        out[f"C{i}"], out[f"a{i}"], out[f"b{i}"] = C, a, b
        out[f"bs{i}"] = np.array(bs, np.int32)
        out[f"perm{i}"], out[f"U{i}"] = np.asarray(F.perm, np.int32), np.triu(F.U)
    return out


def main() -> None:
    from oracle import oracle as O

    O.build()
    with open(os.path.join(HERE, "truths.json"), "w") as f:
        json.dump(truths(), f, indent=1)
    np.savez_compressed(os.path.join(HERE, "sobol_d8.npz"), words=sobol_words())
    np.savez_compressed(os.path.join(HERE, "config0_d16.npz"), **config0(O))
    np.savez_compressed(os.path.join(HERE, "genz_perm.npz"), **genz_perms(O))
    print("wrote", sorted(f for f in os.listdir(HERE) if not f.endswith(".py")))


if __name__ == "__main__":
    main()
