"""Batched uneven-dimension path against the oracle, problem by problem."""
import numpy as np
import pytest

from dbgenzpaper_amd import workloads as W
from dbgenzpaper_amd.batch import Batch
from tests._helpers import rel

pytestmark = [pytest.mark.gpu, pytest.mark.skip(reason="batched create stopped responding on its first GPU run; cause not yet found, path kept off the GPU")]


def _oracle(O, p):
    F = O.prepare_dense(p.Sigma, p.a, p.b)
    bs_i, bs_i2 = O.clamp_block_sizes(p.d, 64, 32)
    return O.qmc_pnorm(F.U, F.a, F.b, m=p.m, n_reps=p.n_reps, shifts=p.shifts, bs_i=bs_i, bs_i2=bs_i2)


def test_batch_fp64_matches_oracle(oracle):
    probs = W.batch_uneven(24, d_min=2, d_max=70, m=600, n_reps=3, seed=31)
    print("batch fp64", [p.d for p in probs], flush=True)
    B = Batch([p.Sigma for p in probs], [p.a for p in probs], [p.b for p in probs],
              [p.shifts for p in probs], T=np.float64, m=600, n_reps=3)
    r = B.run()
    print("tuning", B.tuning(), "ms", B.last_ms(), flush=True)
    for k, p in enumerate(probs):
        ro = _oracle(oracle, p)
        assert r.rank[k] == p.d
        assert rel(r.value[k], ro.value) < 1e-11, (k, p.d, r.value[k], ro.value)
    r2 = B.run()   # graph replay gives the same answer
    assert np.array_equal(r.value, r2.value)


def test_batch_mixed_close_to_fp64():
    probs = W.batch_uneven(32, d_min=2, d_max=256, m=2048, n_reps=4, seed=32)
    args = ([p.Sigma for p in probs], [p.a for p in probs], [p.b for p in probs], [p.shifts for p in probs])
    r32 = Batch(*args, T=np.float32, m=2048, n_reps=4).run()
    r64 = Batch(*args, T=np.float64, m=2048, n_reps=4).run()
    err = np.abs(np.expm1(r32.log_value - r64.log_value))
    print("batch mixed worst rel err", err.max(), flush=True)
    assert np.all(np.isfinite(r32.log_value)) and err.max() < 1e-4
