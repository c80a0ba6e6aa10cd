"""The mixed walk's lower-orthant form (every lower bound -inf: only the upper
bound's tail mass is formed) against the general tail form.  With
a = -1e30 instead of -inf the general form runs, its lower tail mass is
capped at |z| = 30 (about 2^-1300) and enters every state as exactly 0 after
the shift into the upper mass's scale, so both forms give the same estimate
bit for bit on the same factor.  The factor itself can differ: where
erf(b') - erf(a') underflows to 0 during the Genz-Bretz pivoting, the
expected value becomes infinite and a - s is NaN for a = -inf but +inf for
a = -1e30 (src/cholesky.jl's arithmetic, which the device reproduces), so
the deep factor-model problems here are factored with one-column pivot
windows (no pivoting) and the permutations are checked equal first.
Covered: the split-FP16 kernel (d >= 96, with and without the FP64
re-evaluation, Richtmyer and Sobol), the FP32 matrix-core contraction, the
vector-ALU kernel (d < 96), the gradient's record pass and the batch."""
import dataclasses

import numpy as np
import pytest

import dbgenzpaper_amd as dbg
from dbgenzpaper_amd import workloads as W
from dbgenzpaper_amd.batch import Batch

pytestmark = pytest.mark.gpu

A_FAR = -1e30  # finite, so the plan takes the general tail form


def _opts(P):
    # one-column pivot windows: the factor does not depend on the bounds
    return dbg.QMC_opts(np.float32, m=P.m, n_reps=P.n_reps, max_pts=P.m, max_abs_err=0.0,
                        chol_block_size=1, chol_block_size2=1)


def _estimate(P, a, gen="Richtmyer", **kw):
    D = dbg.QMCData(P.Sigma, a, P.b, _opts(P), None, gen, shifts=P.shifts, phi_form="tail", **kw)
    D.loop(0)
    s = D.summary()
    reps = np.array(D.qmc_reps)
    perm = np.array(D.C.perm)
    D.close()
    return s.log_value, reps, perm


@pytest.mark.parametrize("d,gen,kw", [
    (300, "Richtmyer", {}),
    (300, "Richtmyer", {"refine_tiles": 0}),
    (200, "Sobol", {}),
    (260, "Richtmyer", {"contraction": "fp32"}),
    (60, "Richtmyer", {}),
])
def test_lower_orthant_form_equals_general_form(d, gen, kw):
    P = W.factor_model(d, 8, m=4096, n_reps=3, seed=d, gen=gen)
    assert np.all(np.isneginf(P.a))
    print(f"lower orthant d={d} {gen} {kw}", flush=True)
    lv_low, reps_low, perm_low = _estimate(P, P.a, gen, **kw)
    lv_gen, reps_gen, perm_gen = _estimate(P, np.full(d, A_FAR), gen, **kw)
    assert np.array_equal(perm_low, perm_gen)
    assert np.isfinite(lv_low)
    assert lv_low == lv_gen, (lv_low, lv_gen)
    assert np.array_equal(reps_low, reps_gen)


def test_lower_orthant_gradient_record_pass():
    """The gradient's forward with records runs the lower-orthant walk; log P
    is the same bit for bit as with a = -1e30, and the gradient with respect
    to b and Sigma the same up to the order of the FP64 atomic additions that
    reduce the per-workgroup sums (1e-12 relative)."""
    P = W.factor_model(150, 8, m=2048, n_reps=2, seed=5)
    out = []
    for a in (P.a, np.full(P.d, A_FAR)):
        D = dbg.QMCData(P.Sigma, a, P.b, _opts(P), None, "Richtmyer", shifts=P.shifts, phi_form="tail")
        g = D.grad()
        out.append((np.array(g.b), np.array(g.Sigma), g.log_value))
        D.close()
    assert out[0][2] == out[1][2]
    for x, y in ((out[0][0], out[1][0]), (out[0][1], out[1][1])):
        assert np.all(np.isfinite(x))
        assert np.abs(x - y).max() <= 1e-12 * np.abs(y).max(), np.abs(x - y).max()


@pytest.mark.parametrize("width", [16, 32])
def test_lower_orthant_batch(width):
    """A batch mixing lower-orthant problems with problems whose lower bounds
    are finite: each problem's estimate equals the one with a = -1e30 (the
    batch factors in one pivot window; these random-correlation problems with
    b ~ U(-0.5, 2.5) stay far from the underflow that reorders the pivots)."""
    probs = W.batch_uneven(20, d_min=2, d_max=200, m=1024, n_reps=2, seed=41)
    rng = np.random.default_rng(7)
    for k in range(0, len(probs), 3):  # every third problem gets finite lower bounds
        p = probs[k]
        probs[k] = dataclasses.replace(p, a=p.b - rng.uniform(0.5, 3.0, p.d))
    far = [dataclasses.replace(p, a=np.where(np.isneginf(p.a), A_FAR, p.a)) for p in probs]

    def run(ps):
        B = Batch([p.Sigma for p in ps], [p.a for p in ps], [p.b for p in ps], [p.shifts for p in ps],
                  T=np.float32, m=1024, n_reps=2)
        B.set_width(width)  # both batches at the same width (16: vector ALU, 32: matrix cores)
        r = B.run()
        B.close()
        return r

    r_low, r_gen = run(probs), run(far)
    assert np.all(np.isfinite(r_low.log_value))
    assert np.array_equal(r_low.log_value, r_gen.log_value)
