"""Batched uneven-dimension path against the oracle and against single-problem
device solves, problem by problem (no reference counterpart: the reference
runs qmc_pnorm!(QMCData(...)) once per problem, src/qmc.jl:149-202,783-850)."""
import numpy as np
import pytest

import dbgenzpaper_amd as dbg
from dbgenzpaper_amd import workloads as W
from dbgenzpaper_amd import _lib as L
from dbgenzpaper_amd.batch import Batch
from tests._helpers import rel

pytestmark = pytest.mark.gpu


def _oracle(O, p):
    F = O.prepare_dense(p.Sigma, p.a, p.b)
    bs_i, bs_i2 = O.clamp_block_sizes(p.d, 64, 32)
    return O.qmc_pnorm(F.U, F.a, F.b, m=p.m, n_reps=p.n_reps, shifts=p.shifts, bs_i=bs_i, bs_i2=bs_i2)


def _args(probs):
    return ([p.Sigma for p in probs], [p.a for p in probs], [p.b for p in probs], [p.shifts for p in probs])


def test_batch_fp64_matches_oracle(oracle):
    probs = W.batch_uneven(24, d_min=2, d_max=70, m=600, n_reps=3, seed=31)
    print("batch fp64 K=24 m=600 dims", [p.d for p in probs], flush=True)
    B = Batch(*_args(probs), T=np.float64, m=600, n_reps=3)
    r = B.run()
    print("tuning", B.tuning(), "ms", B.last_ms(), flush=True)
    for k, p in enumerate(probs):
        ro = _oracle(oracle, p)
        assert r.rank[k] == p.d
        assert rel(r.value[k], ro.value) < 1e-11, (k, p.d, r.value[k], ro.value)
        assert abs(r.err3[k] - ro.err3) <= 1e-9 * max(ro.err3, 1e-300) + 1e-15, (k, r.err3[k], ro.err3)
    r2 = B.run()  # graph replay gives the same answer
    assert np.array_equal(r.value, r2.value)
    B.close()


def test_batch_fp64_matches_single_plans():
    """Each problem of the batch equals the single-problem FP64 plan on the
    same inputs and shifts (erf form, Richtmyer, one round)."""
    probs = W.batch_uneven(12, d_min=2, d_max=256, m=1000, n_reps=3, seed=33)
    print("batch vs plans K=12 m=1000 dims", [p.d for p in probs], flush=True)
    r = Batch(*_args(probs), T=np.float64, m=1000, n_reps=3).run()
    for k, p in enumerate(probs):
        opts = dbg.QMC_opts(np.float64, m=p.m, n_reps=p.n_reps, max_pts=p.m, max_abs_err=0.0)
        D = dbg.QMCData(p.Sigma, p.a, p.b, opts, None, "Richtmyer", shifts=p.shifts, phi_form="erf")
        v, e, n = dbg.qmc_pnorm(D)
        D.close()
        assert rel(r.value[k], v) < 1e-12, (k, p.d, r.value[k], v)


@pytest.mark.parametrize("width", [16, 32])
def test_batch_mixed_close_to_fp64(width):
    """The mixed batch at a fixed block width (16: vector-ALU contraction;
    32: FP16 matrix cores) against the FP64 batch."""
    probs = W.batch_uneven(32, d_min=2, d_max=256, m=2048, n_reps=4, seed=32)
    print(f"batch mixed K=32 m=2048 width {width} dims", sorted(p.d for p in probs), flush=True)
    B32 = Batch(*_args(probs), T=np.float32, m=2048, n_reps=4)
    B32.set_width(width)
    assert B32.tuning()[0] == width
    r32 = B32.run()
    B32.close()
    r64 = Batch(*_args(probs), T=np.float64, m=2048, n_reps=4).run()
    err = np.abs(np.expm1(r32.log_value - r64.log_value))
    print("batch mixed worst rel err", err.max(), flush=True)
    assert np.all(np.isfinite(r32.log_value)) and err.max() < 1e-4


def test_batch_matrix_core_block_edges(oracle):
    """The matrix-core contraction (mixed, width 32) on dimensions around the
    32-wide blocks (d - 1 = 31, 32, 33, 63, 64, 65, 96, 128, 255), one and two
    points past a tile, against the oracle in FP64 on the same factor inputs
    (the mixed path's bound, 1e-4 relative) and against the vector-ALU width
    16 (the same walk on differently rounded sums)."""
    rng = np.random.default_rng(36)
    probs = []
    for d in (2, 32, 33, 34, 64, 65, 66, 97, 129, 256):
        P = W.dense_random(d, m=513, n_reps=2, seed=int(rng.integers(1 << 30)))
        P.shifts = np.asfortranarray(rng.random((d - 1, 2)))
        probs.append(P)
    out = {}
    for width in (16, 32):
        B = Batch(*_args(probs), T=np.float32, m=513, n_reps=2)
        B.set_width(width)
        out[width] = B.run()
        r2 = B.run()
        assert np.array_equal(out[width].value, r2.value)  # graph replay
        B.close()
    for k, p in enumerate(probs):
        ro = _oracle(oracle, p)
        for width in (16, 32):
            assert rel(out[width].value[k], ro.value) < 1e-4, (width, p.d, out[width].value[k], ro.value)
        assert rel(out[32].value[k], out[16].value[k]) < 1e-4


def test_batch_set_width_rejects():
    probs = W.batch_uneven(2, d_min=3, d_max=5, m=256, n_reps=1, seed=37)
    B = Batch(*_args(probs), T=np.float32, m=256, n_reps=1)
    with pytest.raises(L.DbgError):
        B.set_width(12)
    B.close()


def test_batch_ragged_edges(oracle):
    """d = 1 (no sampled dimension), d = 2, the largest block-width remainder,
    and m not a multiple of the 256-point tile, in one batch."""
    rng = np.random.default_rng(34)
    probs = []
    for d in (1, 2, 33, 2, 1, 97):
        P = W.dense_random(d, m=300, n_reps=2, seed=int(rng.integers(1 << 30)))
        P.shifts = np.asfortranarray(rng.random((d - 1, 2)))
        probs.append(P)
    print("batch ragged dims", [p.d for p in probs], flush=True)
    r = Batch(*_args(probs), T=np.float64, m=300, n_reps=2).run()
    for k, p in enumerate(probs):
        if p.d == 1:
            from scipy.stats import norm
            assert rel(r.value[k], norm.cdf(p.b[0]) - norm.cdf(p.a[0])) < 1e-14
            continue
        ro = _oracle(oracle, p)
        assert rel(r.value[k], ro.value) < 1e-11, (k, p.d, r.value[k], ro.value)


def test_batch_large_d_repeatable():
    """48 problems of d = 200..256 (long copy loops in the batched
    constructor): every factor of full rank, replays identical, the FP64
    batch equal to single-problem plans and the mixed batch finite and close
    to it."""
    probs = W.batch_uneven(48, d_min=200, d_max=256, m=512, n_reps=2, seed=35)
    print("batch large d K=48 m=512 dims", sorted(p.d for p in probs), flush=True)
    B64 = Batch(*_args(probs), T=np.float64, m=512, n_reps=2)
    r1, r2 = B64.run(), B64.run()
    B64.close()
    assert np.array_equal(r1.rank, [p.d for p in probs])
    assert np.array_equal(r1.value, r2.value)
    for k in range(0, 48, 6):
        p = probs[k]
        opts = dbg.QMC_opts(np.float64, m=p.m, n_reps=p.n_reps, max_pts=p.m, max_abs_err=0.0)
        D = dbg.QMCData(p.Sigma, p.a, p.b, opts, None, "Richtmyer", shifts=p.shifts, phi_form="erf")
        v, _, _ = dbg.qmc_pnorm(D)
        D.close()
        assert rel(r1.value[k], v) < 1e-12, (k, p.d, r1.value[k], v)
    r32 = Batch(*_args(probs), T=np.float32, m=512, n_reps=2).run()
    assert np.array_equal(r32.rank, r1.rank)
    err = np.abs(np.expm1(r32.log_value - r1.log_value))
    print("  mixed worst rel err", err.max(), flush=True)
    assert np.all(np.isfinite(r32.log_value)) and err.max() < 1e-4
