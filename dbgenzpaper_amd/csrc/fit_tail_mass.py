"""Fit of the polynomial P in tail_mass_f32 (tailq.h).

erfc(z) = t exp(-z^2 + P(t)) with t = 2 / (2 + z) defines P on t in (0, 1]
(z = 2 (1/t - 1) >= 0); P(0) = -log(2 sqrt(pi)) is its limit as z -> inf.
P is fitted as a polynomial in u = 2t - 1 by least squares on 600 Chebyshev
nodes of t against 50-digit values (mpmath), and checked on a uniform grid of
6001 points.  Prints the maximum fit error and the coefficients, highest power
first, rounded to FP32.  Run: python fit_tail_mass.py [degree]  (default 11).
"""
import sys

import mpmath as mp
import numpy as np

mp.mp.dps = 50


def target(t):
    t = mp.mpf(t)
    if t == 0:
        return -mp.log(2 * mp.sqrt(mp.pi))
    z = 2 * (1 / t - 1)
    return mp.log(mp.erfc(z)) + z * z - mp.log(t)


def main(deg=11, nodes=600):
    x = 0.5 - 0.5 * np.cos(np.pi * (np.arange(nodes) + 0.5) / nodes)
    y = np.array([float(target(xi)) for xi in x])
    c, *_ = np.linalg.lstsq(np.vander(2 * x - 1, deg + 1), y, rcond=None)
    xs = np.linspace(0, 1, 6001)
    ys = np.array([float(target(xi)) for xi in xs])
    err = np.abs(np.polyval(c, 2 * xs - 1) - ys).max()
    print(f"degree {deg}: max fit error {err:.3e}")
    print(", ".join("%.9ef" % np.float32(ci) for ci in c))
    return c


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 11)
