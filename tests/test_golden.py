"""The committed golden fixtures (tests/golden/, written by make_golden.py)
against the CPU oracle and the host-side generators (no GPU)."""
import json
import os

import numpy as np
import pytest

from dbgenzpaper_amd.sobol import direction_integers

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def truths():
    with open(os.path.join(GOLD, "truths.json")) as f:
        return json.load(f)


def test_config0_fixture_reproduced(oracle):
    """d = 16 headline CPU case: the oracle's factor, permutation and
    estimate are unchanged (bitwise for the permutation, 1e-15 otherwise)."""
    g = load("config0_d16.npz")
    F = oracle.prepare_dense(g["Sigma"], g["a"], g["b"])
    assert np.array_equal(F.perm, g["perm"])
    assert np.allclose(F.U, g["U"], rtol=0, atol=1e-15)
    bs_i, bs_i2 = oracle.clamp_block_sizes(16, 2 ** 10, 2 ** 6)
    r = oracle.qmc_pnorm(F.U, F.a, F.b, m=int(g["m"]), n_reps=int(g["n_reps"]), shifts=g["shifts"],
                         bs_i=bs_i, bs_i2=bs_i2)
    assert abs(r.value / float(g["value"]) - 1) < 1e-15
    assert abs(r.err3 / float(g["err3"]) - 1) < 1e-12


def test_genz_perm_fixtures_reproduced(oracle):
    g = load("genz_perm.npz")
    for i in range(4):
        F = oracle.cholesky_genz(g[f"C{i}"], g[f"a{i}"], g[f"b{i}"], *[int(v) for v in g[f"bs{i}"]])
        assert np.array_equal(F.perm, g[f"perm{i}"]), i
        assert np.allclose(np.triu(F.U), g[f"U{i}"], rtol=0, atol=1e-14), i


def test_sobol_words_fixture(oracle):
    """The first 1024 unscrambled Sobol points of 8 dimensions (SciPy, Joe-Kuo
    directions) from the oracle's Gray-code generator: point i is sequence
    index i (SciPy's point 0 is the origin)."""
    w = load("sobol_d8.npz")["words"]
    X = oracle.sobol_block(direction_integers(8), 0, 1023)
    assert np.array_equal(X.astype(np.uint32), w[1:])
    assert not w[0].any()


def test_truths_fixture(oracle):
    t = truths()
    # Sigma = I: every weight is the closed-form product
    idt = t["identity_d16"]
    a, b = np.array(idt["a"]), np.array(idt["b"])
    F = oracle.prepare_dense(np.eye(16), a, b)
    sh = np.random.default_rng(0).random((15, 4))
    r = oracle.qmc_pnorm(F.U, F.a, F.b, m=1024, n_reps=4, shifts=sh, bs_i=15, bs_i2=15)
    assert abs(r.value / idt["value"] - 1) < 1e-13
    # d = 2 against the bivariate normal CDF, inside 3 SE
    bv = t["bivariate"]
    S = np.array(bv["Sigma"])
    F = oracle.prepare_dense(S, np.array([-np.inf, -np.inf]), np.array(bv["b"]))
    sh = np.random.default_rng(1).random((1, 8))
    r = oracle.qmc_pnorm(F.U, F.a, F.b, m=2 ** 14, n_reps=8, shifts=sh, bs_i=1, bs_i2=1)
    assert abs(r.value - bv["value"]) < max(3 * r.err3, 1e-6)
    # the published fixed_dense medians agree with the quadrature truth family
    eq = t["equicorrelated"]
    assert abs(eq["256"]["value"] - t["published_fixed_dense_medians"]["256"]) < 1e-3


def test_quantile_fixture_fp32(tmp_path):
    """The mixed path's FP32 tail quantile against the 40-digit fixture."""
    from tests.test_host import _host_build
    import ctypes

    lib = _host_build(tmp_path, '#include "tailq.h"\n'
                      'extern "C" float tq(float t, int k) { return dbg::tail_quantile_f32(t, k); }\n', "tqg")
    lib.tq.restype = ctypes.c_float
    lib.tq.argtypes = [ctypes.c_float, ctypes.c_int]
    worst = 0.0
    for tt, k, y in truths()["quantiles_t_k_y"]:
        worst = max(worst, abs(lib.tq(tt, int(k)) - y) / max(abs(y), 1.0))
    assert worst < 4 * np.finfo(np.float32).eps, worst


@pytest.mark.parametrize("d", ["16", "64"])
def test_equicorrelated_fixture(oracle, d):
    from dbgenzpaper_amd import workloads as W

    e = truths()["equicorrelated"][d]
    n = int(d)
    P = W.equicorrelated(n, m=2 ** 12, n_reps=12)
    assert np.array_equal(P.a, np.array(e["a"])) and np.array_equal(P.b, np.array(e["b"]))
    F = oracle.prepare_dense(P.Sigma, P.a, P.b)
    bs_i, bs_i2 = oracle.clamp_block_sizes(n, 2 ** 10, 2 ** 6)
    r = oracle.qmc_pnorm(F.U, F.a, F.b, m=P.m, n_reps=P.n_reps, shifts=P.shifts, bs_i=bs_i, bs_i2=bs_i2)
    assert abs(r.value - e["value"]) < max(2 * r.err3, 1e-4)
