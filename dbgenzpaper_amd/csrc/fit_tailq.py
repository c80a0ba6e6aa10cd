"""Fits of the polynomials in tailq.h (tail_mass_f32, tail_quantile_f32).

mass:      erfc(z) = t exp(-z^2 + P(t)) with t = 2 / (2 + z) defines P on t in
           (0, 1] (z = 2 (1/t - 1) >= 0); P(0) = -log(2 sqrt(pi)) is its limit
           as z -> inf.  P is fitted as a polynomial in u = 2t - 1 by least
           squares on 600 Chebyshev nodes of t, checked on 6001 uniform points.
quantile:  y = Phi^{-1}(p) for p <= 1/2 as a function of w = -log(4p(1 - p)),
           in two regions, each fitted for least relative error on 500
           Chebyshev nodes:  y / (2p - 1) as a polynomial in w - 3.125 for
           w < 6.25;  -y / s as a polynomial in 1/s - 0.21, s = sqrt(w), for
           s in [2.5, 50] (p from 4.8e-4 down to about 2^-3600).
Values come from mpmath at 50-60 digits.  Prints the maximum relative fit
error and the coefficients, highest power first, rounded to FP32.
Run: python fit_tailq.py mass | quantile
"""
import sys

import mpmath as mp
import numpy as np

mp.mp.dps = 60


def cheb(a, b, n):
    return (a + b) / 2 + (b - a) / 2 * np.cos(np.pi * (np.arange(n) + 0.5) / n)


def show(name, c, err, kind="relative"):
    print(f"{name}: max {kind} fit error {err:.3e}")
    print(", ".join("%.9ef" % np.float32(ci) for ci in c))


def fit_mass(deg=11, nodes=600):
    def target(t):
        t = mp.mpf(t)
        if t == 0:
            return -mp.log(2 * mp.sqrt(mp.pi))
        z = 2 * (1 / t - 1)
        return mp.log(mp.erfc(z)) + z * z - mp.log(t)

    x = 0.5 - 0.5 * np.cos(np.pi * (np.arange(nodes) + 0.5) / nodes)
    y = np.array([float(target(xi)) for xi in x])
    c, *_ = np.linalg.lstsq(np.vander(2 * x - 1, deg + 1), y, rcond=None)
    xs = np.linspace(0, 1, 6001)
    ys = np.array([float(target(xi)) for xi in xs])
    show("P", c, np.abs(np.polyval(c, 2 * xs - 1) - ys).max(), "absolute")


def quantile_of_w(w):
    """Phi^{-1}(p) and p for the p <= 1/2 with 4 p (1 - p) = exp(-w)."""
    w = mp.mpf(w)
    e = mp.exp(-w)
    p = e / (2 * (1 + mp.sqrt(1 - e)))
    if w < 50:
        return mp.sqrt(2) * mp.erfinv(2 * p - 1), p
    logp, y = mp.log(p), -mp.sqrt(2 * w)
    for _ in range(200):  # Newton on log Phi(y) = log p
        yn = y - (mp.log(mp.ncdf(y)) - logp) * mp.ncdf(y) / mp.npdf(y)
        if abs(yn - y) < mp.mpf(10) ** -45:
            break
        y = yn
    return yn, p


def relfit(x, y, deg):
    c, *_ = np.linalg.lstsq(np.vander(x, deg + 1) / y[:, None], np.ones_like(y), rcond=None)
    return c, np.abs(np.polyval(c, x) / y - 1).max()


def fit_quantile(nodes=500):
    ws = cheb(0.0, 6.25, nodes)
    f1 = []
    for w in ws:
        y, p = quantile_of_w(w)
        f1.append(float(y / (2 * p - 1)))
    show("C1, region w < 6.25", *relfit(ws - 3.125, np.array(f1), 10))
    s2 = cheb(2.5, 50.0, nodes)
    g2 = np.array([float(-quantile_of_w(s * s)[0] / s) for s in s2])
    show("CB, region s in [2.5, 50]", *relfit(1 / s2 - 0.21, g2, 12))

if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "mass"
    fit_mass() if which == "mass" else fit_quantile()
