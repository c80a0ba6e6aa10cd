"""FP64 re-evaluation of the heaviest tiles of a mixed round (no reference
counterpart).  With every tile re-evaluated the mixed plan must reproduce
the plain FP64 plan (same points, shifts and tail form) to rounding; with a
few tiles, its distance from the FP64 plan is bounded by the share of the sum
left in FP32 times the FP32 walk's own relative error."""
import numpy as np
import pytest

from dbgenzpaper_amd import workloads as W

pytestmark = pytest.mark.gpu


def _pair(dbg, P, gen, refine):
    o32 = dbg.QMC_opts(np.float32, m=P.m, n_reps=P.n_reps, max_pts=P.m, max_abs_err=0.0)
    o64 = dbg.QMC_opts(np.float64, m=P.m, n_reps=P.n_reps, max_pts=P.m, max_abs_err=0.0)
    Dm = dbg.QMCData(P.Sigma, P.a, P.b, o32, None, gen, shifts=P.shifts, refine_tiles=refine)
    D64 = dbg.QMCData(P.Sigma, P.a, P.b, o64, None, gen, shifts=P.shifts, phi_form="tail")
    dbg.qmc_pnorm(Dm)
    dbg.qmc_pnorm(D64)
    out = (Dm.last_result.log_value, D64.last_result.log_value, Dm.refine_info(),
           np.array(Dm.qmc_reps), np.array(D64.qmc_reps))
    Dm.close()
    D64.close()
    return out


@pytest.mark.parametrize("d,family,gen,m", [(40, "random", "Richtmyer", 1000), (300, "factor", "Richtmyer", 2048),
                                            (300, "factor", "Sobol", 1000), (150, "random", "Richtmyer", 2304)])
def test_refine_all_tiles_equals_fp64(dbg, d, family, gen, m):
    """refine_tiles >= tiles per shift: every mixed tile is replaced by its
    FP64 tiles (the split kernel and the FP32 matrix-core kernel have 256-
    sample tiles against 128 of the FP64 kernel; the vector kernel of small d
    has the same tile), so the mixed estimate is the FP64 estimate.  m not a
    multiple of 256 leaves the last FP64 tile of a shift empty."""
    P = (W.dense_random(d, m=m, n_reps=3, seed=d, gen=gen) if family == "random"
         else W.factor_model(d, m=m, n_reps=3, seed=d, gen=gen))
    print(f"refine all d={d} {family} {gen} m={m}", flush=True)
    lm, l64, (k, share, ms), rm, r64 = _pair(dbg, P, gen, refine=1 << 20)
    print(f"  log P mixed+refine {lm:.15f} fp64 {l64:.15f} tiles {k} share {share:.2e} {ms:.3f} ms", flush=True)
    assert k == (m + 255) // 256 if d >= 96 else k >= 1
    assert share == 0.0
    assert abs(lm - l64) <= 1e-12 * max(1.0, abs(l64))
    assert np.allclose(rm, r64, rtol=1e-12, atol=0)


@pytest.mark.parametrize("d", [500, 1100])
def test_refine_heaviest_tiles(dbg, d):
    """The deep lower orthant of the factor model, where a few samples hold
    the estimate: refining a few tiles per shift brings the mixed estimate to
    the FP64 one within share x 1e-2 (1e-2: a bound on the FP32 walk's
    per-sample relative error at these shapes, which is about 1e-3), and well
    below the unrefined mixed error."""
    P = W.factor_model(d, 8, m=4096, n_reps=4, seed=d)
    print(f"refine heaviest d={d}", flush=True)
    l0, l64, _, _, _ = _pair(dbg, P, "Richtmyer", refine=0)
    lm, _, (k, share, ms), _, _ = _pair(dbg, P, "Richtmyer", refine=4)
    print(f"  |dlogP| unrefined {abs(l0 - l64):.2e} refined {abs(lm - l64):.2e} share {share:.2e} "
          f"{ms:.3f} ms", flush=True)
    assert k == 4 and 0.0 <= share <= 1.0
    assert abs(lm - l64) <= share * 1e-2 + 1e-12 * abs(l64)
