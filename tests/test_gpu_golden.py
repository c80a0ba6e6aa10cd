"""The device path against the committed golden fixtures (tests/golden/)."""
import numpy as np
import pytest

from tests.test_golden import load, truths

pytestmark = pytest.mark.gpu


def test_genz_perm_fixtures(dbg):
    g = load("genz_perm.npz")
    for i in range(4):
        bs = [int(v) for v in g[f"bs{i}"]]
        print(f"golden genz case {i} n={g[f'C{i}'].shape[0]} bs={bs}", flush=True)
        F = dbg.cholesky_genz(np.asfortranarray(g[f"C{i}"].copy()), g[f"a{i}"].copy(), g[f"b{i}"].copy(), *bs)
        assert np.array_equal(F.perm, g[f"perm{i}"]), i
        assert np.abs(np.triu(F.U) - g[f"U{i}"]).max() < 1e-12, i


def test_config0_fixture(dbg):
    """d = 16, N = 2^14 x 8, FP64: the device estimate equals the fixture's
    oracle estimate to 1e-12 relative."""
    g = load("config0_d16.npz")
    print("golden config0", flush=True)
    opts = dbg.QMC_opts(np.float64, m=int(g["m"]), n_reps=int(g["n_reps"]), max_pts=int(g["m"]), max_abs_err=0.0)
    D = dbg.QMCData(g["Sigma"], g["a"], g["b"], opts, None, "Richtmyer", shifts=g["shifts"])
    assert np.array_equal(D.C.perm, g["perm"])
    v, e, n = dbg.qmc_pnorm(D)
    assert abs(v / float(g["value"]) - 1) < 1e-12
    assert abs(e / float(g["err3"]) - 1) < 1e-8
    assert n == int(g["m"])


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_identity_fixture(dbg, T):
    t = truths()["identity_d16"]
    print(f"golden identity {np.dtype(T).name}", flush=True)
    opts = dbg.QMC_opts(T, m=1024, n_reps=4, max_pts=1024, max_abs_err=0.0)
    sh = np.random.default_rng(0).random((15, 4))
    # refine_tiles=0: the mixed estimate comes from the FP32 walk alone
    D = dbg.QMCData(np.eye(16), np.array(t["a"]), np.array(t["b"]), opts, None, "Richtmyer", shifts=sh,
                    refine_tiles=0)
    v, e, _ = dbg.qmc_pnorm(D)
    tol = 1e-13 if T == np.float64 else 2e-6
    assert abs(v / t["value"] - 1) < tol


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_equicorrelated_fixture(dbg, T):
    from dbgenzpaper_amd import workloads as W

    e = truths()["equicorrelated"]["256"]
    print(f"golden equicorrelated d=256 {np.dtype(T).name}", flush=True)
    P = W.equicorrelated(256, m=2 ** 14, n_reps=12)
    opts = dbg.QMC_opts(T, m=P.m, n_reps=P.n_reps, max_pts=P.m, max_abs_err=0.0)
    D = dbg.QMCData(P.Sigma, P.a, P.b, opts, None, "Richtmyer", shifts=P.shifts)
    v, err, _ = dbg.qmc_pnorm(D)
    assert abs(v - e["value"]) < max(2 * err, 1e-4)
