"""Sparse / banded integration on the device against the oracle (same
unpivoted factor, points and shifts)."""
import numpy as np
import pytest

from dbgenzpaper_amd import workloads as W
from tests._helpers import oracle_same_inputs, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("gen", ["Richtmyer", "Sobol"])
def test_banded_fp64(dbg, oracle, gen):
    P = W.banded(200, 5, m=1500, n_reps=3, gen=gen)
    opts = dbg.QMC_opts(np.float64, m=P.m, n_reps=P.n_reps, max_pts=P.m)
    print(f"banded {gen}", flush=True)
    D = dbg.QMCDataSparse(P.Sigma, P.a, P.b, opts, None, gen, shifts=P.shifts)
    v, _, _ = dbg.qmc_pnorm(D)
    r = oracle_same_inputs(oracle, D)
    assert rel(v, r.value) < 1e-12


def test_dense_equals_sparse_on_same_factor(dbg, oracle):
    """Invariant: the dense loop over the unpivoted factor equals the sparse
    loop (data/sims/sparse_dense_vals.csv fixed_dense rows agree to 1e-15)."""
    P = W.banded(60, 4, m=2048, n_reps=2)
    opts = dbg.QMC_opts(np.float64, m=P.m, n_reps=P.n_reps, max_pts=P.m)
    Ds = dbg.QMCDataSparse(P.Sigma, P.a, P.b, opts, None, "Richtmyer", shifts=P.shifts)
    vs, _, _ = dbg.qmc_pnorm(Ds)
    r = oracle_same_inputs(oracle, Ds, kind="dense")
    assert rel(vs, r.value) < 1e-12


def test_reference_sparse_testset(dbg):
    """test/runtests.jl:175-201: tridiagonal 1.8/0.25, b = 1.5; loose check
    against scipy's Genz-based MVN CDF (the reference uses MvNormalCDF)."""
    from scipy.stats import multivariate_normal

    n = 8
    S = 1.8 * np.eye(n) + 0.25 * (np.eye(n, k=1) + np.eye(n, k=-1))
    a, b = np.full(n, -np.inf), np.full(n, 1.5)
    opts = dbg.QMC_opts(np.float64, m=2 ** 11, max_pts=2 ** 11 * 4, max_abs_err=0.0,
                        block_size_i=2 ** 6, block_size_i2=2 ** 5, block_size_j=2 ** 6)
    v, e, _ = dbg.qmc_pnorm(dbg.QMCDataSparse(S, a, b, opts, 21, "Richtmyer"))
    ref = multivariate_normal(np.zeros(n), S).cdf(b)
    assert np.isfinite(v) and np.isfinite(e) and abs(v - ref) < 0.06


def test_sparseqmc_module(dbg, oracle):
    P = W.banded(40, 3, m=1024, n_reps=3, gen="Sobol")
    opts = dbg.SparseQMC.QMC_opts(np.float64, m=P.m, n_reps=P.n_reps, max_pts=P.m * 2, max_abs_err=0.0)
    D = dbg.SparseQMC.QMCDataSparse(P.Sigma, P.a, P.b, opts=opts)
    v, _, npts = dbg.SparseQMC.qmc_pnorm(D, shifts=P.shifts)
    r = oracle_same_inputs(oracle, D, kind="sparseqmc", gen_state=1)
    assert npts == 2 * P.m and rel(v, r.value) < 1e-12


def test_banded_mixed(dbg, oracle):
    P = W.banded(300, 8, m=4096, n_reps=3)
    o32 = dbg.QMC_opts(np.float32, m=P.m, n_reps=P.n_reps, max_pts=P.m)
    o64 = dbg.QMC_opts(np.float64, m=P.m, n_reps=P.n_reps, max_pts=P.m)
    v, _, _ = dbg.qmc_pnorm(dbg.QMCDataSparse(P.Sigma, P.a, P.b, o32, None, "Richtmyer", shifts=P.shifts))
    D64 = dbg.QMCDataSparse(P.Sigma, P.a, P.b, o64, None, "Richtmyer", shifts=P.shifts)
    r = oracle_same_inputs(oracle, D64)
    print("banded mixed rel", rel(v, r.value), flush=True)
    assert rel(v, r.value) < 2e-5


@pytest.mark.parametrize("bw", [1, 20, 40, 90])
def test_banded_row_layouts(dbg, oracle, bw):
    """Each place the sparse kernel keeps earlier y rows: band reach 0 and 19
    (LDS rings of 16 and 32 rows), 39 (a 64-row ring in FP32, the global
    workspace in FP64) and 89 (the global workspace in both)."""
    P = W.banded(160, bw, m=1024, n_reps=2, seed=bw)
    print(f"banded layouts bw={bw}", flush=True)
    o64 = dbg.QMC_opts(np.float64, m=P.m, n_reps=P.n_reps, max_pts=P.m)
    o32 = dbg.QMC_opts(np.float32, m=P.m, n_reps=P.n_reps, max_pts=P.m)
    D64 = dbg.QMCDataSparse(P.Sigma, P.a, P.b, o64, None, "Richtmyer", shifts=P.shifts)
    v64, _, _ = dbg.qmc_pnorm(D64)
    v32, _, _ = dbg.qmc_pnorm(dbg.QMCDataSparse(P.Sigma, P.a, P.b, o32, None, "Richtmyer", shifts=P.shifts))
    r = oracle_same_inputs(oracle, D64)
    assert rel(v64, r.value) < 1e-12
    assert rel(v32, r.value) < 1e-4
