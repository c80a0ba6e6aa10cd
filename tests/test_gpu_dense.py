"""Dense integration on the device against the oracle on the same factor,
points and shifts (FP64 path: reference arithmetic; mixed path: within a
stated tolerance)."""
import numpy as np
import pytest
from scipy.special import ndtr

from dbgenzpaper_amd import workloads as W
from tests._helpers import oracle_same_inputs, rel

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12      # relative, same factor / points / shifts


def _plan(dbg, P, T=np.float64, gen="Richtmyer", phi_form=None, refine_tiles=None, **ow):
    kw = dict(m=P.m, n_reps=P.n_reps, max_pts=P.m, max_abs_err=0.0)
    kw.update(ow)
    opts = dbg.QMC_opts(T, **kw)
    return dbg.QMCData(P.Sigma, P.a, P.b, opts, None, gen, shifts=P.shifts, phi_form=phi_form,
                       refine_tiles=refine_tiles)


def test_config0_fp64(dbg, oracle):
    """d = 16 random correlation, a = -inf, b ~ U(0, 2), N = 2^14 x 8."""
    P = W.dense_random(16)
    print("config0 fp64", flush=True)
    D = _plan(dbg, P)
    v, e, npts = dbg.qmc_pnorm(D)
    r = oracle_same_inputs(oracle, D)
    assert npts == P.m
    assert rel(v, r.value) < F64_TOL, (v, r.value)
    assert rel(e, r.err3) < 1e-8
    assert np.allclose(D.qmc_reps, r.rep_means, rtol=F64_TOL, atol=0)


@pytest.mark.parametrize("d,m,R", [(2, 1000, 3), (3, 257, 2), (17, 999, 4), (33, 2048, 2), (100, 700, 3)])
def test_ragged_shapes_fp64(dbg, oracle, d, m, R):
    P = W.dense_random(d, m=m, n_reps=R, seed=d)
    print(f"ragged d={d} m={m} R={R}", flush=True)
    D = _plan(dbg, P)
    v, _, _ = dbg.qmc_pnorm(D)
    r = oracle_same_inputs(oracle, D)
    assert rel(v, r.value) < F64_TOL


def test_erfc_form(dbg, oracle):
    P = W.dense_random(24, m=4096, n_reps=4, seed=11)
    D = _plan(dbg, P)
    De = dbg.QMCData_Distributions(D)
    v, _, _ = dbg.qmc_pnorm(De)
    r = oracle_same_inputs(oracle, De)
    assert rel(v, r.value) < F64_TOL
    v0, _, _ = dbg.qmc_pnorm(D)
    assert abs(v - v0) < 4e-14   # erf vs erfc form on the same points (data/sims/table_precision.tex)


def test_sobol(dbg, oracle):
    P = W.dense_random(20, m=2048, n_reps=3, seed=4, gen="Sobol")
    D = _plan(dbg, P, gen="Sobol")
    v, _, _ = dbg.qmc_pnorm(D)
    r = oracle_same_inputs(oracle, D)
    assert rel(v, r.value) < F64_TOL


def test_adaptive_rounds_and_state(dbg, oracle):
    """max_abs_err = 0 forces max_pts/m rounds (test/runtests.jl:157-167);
    the generator state carries over to the next call like the reference's."""
    P = W.dense_random(8, m=2 ** 11, n_reps=12, seed=42)
    D = _plan(dbg, P, max_pts=2 ** 11 * 4)
    v, e, npts = dbg.qmc_pnorm(D)
    assert npts == 2 ** 13 and D.last_result.rounds == 4
    r = oracle_same_inputs(oracle, D, gen_state=0)
    assert r.rounds == 4 and rel(v, r.value) < F64_TOL and rel(e, r.err3) < 1e-8
    assert D.gen_state == 3
    v2, _, _ = dbg.qmc_pnorm(D)
    r2 = oracle_same_inputs(oracle, D, gen_state=3)
    assert rel(v2, r2.value) < F64_TOL and D.gen_state == 6


def test_identity_closed_form(dbg):
    n = 12
    rng = np.random.default_rng(3)
    b = rng.uniform(0.2, 2, n)
    a = -rng.uniform(0.2, 2, n)
    opts = dbg.QMC_opts(np.float64, m=4096, n_reps=4, max_pts=4096)
    v, e, _ = dbg.qmc_pnorm(dbg.QMCData(np.eye(n), a, b, opts, 1))
    exact = np.prod(ndtr(b) - ndtr(a))
    assert rel(v, exact) < 1e-13 and e < 1e-15


@pytest.mark.parametrize("d", [100, 500])
def test_underflow_tail_form(dbg, oracle, d):
    """Lower orthant of a factor model: P is far below 1e-308 and every weight
    of the reference forms is 0 (per-dimension masses fall under the [-9, 9]
    clamp).  The tail form keeps the weights; the log of the estimate matches
    the oracle's tail form on the same points."""
    P = W.factor_model(d, m=1024, n_reps=3)
    D = _plan(dbg, P, phi_form="tail")
    dbg.qmc_pnorm(D)
    r = oracle_same_inputs(oracle, D)
    lv = D.last_result.log_value
    print(f"d={d} log P {lv} oracle {r.log_value}", flush=True)
    assert np.isfinite(lv) and abs(lv - r.log_value) < 1e-9 * abs(r.log_value)


@pytest.mark.parametrize("d,form", [(24, "tail"), (150, "erf"), (150, "tail"), (300, "erfc")])
def test_forms_fp64(dbg, oracle, d, form):
    """FP64 forms on both dense kernels (d >= 96 runs the matrix-core kernel)."""
    P = W.dense_random(d, m=1500, n_reps=2, seed=d)
    D = _plan(dbg, P, phi_form=form)
    v, _, _ = dbg.qmc_pnorm(D)
    r = oracle_same_inputs(oracle, D)
    print(f"d={d} {form} rel {rel(v, r.value):.2e}", flush=True)
    assert rel(v, r.value) < F64_TOL


@pytest.mark.parametrize("d,gen", [(150, "Richtmyer"), (400, "Sobol")])
def test_mixed_mfma(dbg, oracle, d, gen):
    """FP32 mixed on the matrix-core kernel against the FP64 oracle (tail
    form, same factor, points and shifts); no tile re-evaluated in FP64."""
    P = W.dense_random(d, m=4096, n_reps=3, seed=d, gen=gen)
    D = _plan(dbg, P, T=np.float32, gen=gen, refine_tiles=0)
    v, _, _ = dbg.qmc_pnorm(D)
    D64 = _plan(dbg, P, gen=gen, phi_form="tail")
    r = oracle_same_inputs(oracle, D64)
    print(f"mixed mfma d={d} {gen} rel err {rel(v, r.value):.3e} (se {r.err3 / 3 / r.value:.3e})", flush=True)
    assert rel(v, r.value) < 2e-5  # measured 1.2e-6 (d = 150), 3.8e-6 (d = 400)


@pytest.mark.parametrize("gen", ["Richtmyer", "Sobol"])
def test_mixed_within_tolerance(dbg, oracle, gen):
    P = W.dense_random(64, m=8192, n_reps=4, seed=21, gen=gen)
    D = _plan(dbg, P, T=np.float32, gen=gen, refine_tiles=0)
    v, e, _ = dbg.qmc_pnorm(D)
    # FP64 oracle on the same points, shifts and (FP64) factor of an FP64 plan
    D64 = _plan(dbg, P, gen=gen)
    r = oracle_same_inputs(oracle, D64)
    print(f"mixed {gen} rel err {rel(v, r.value):.3e} (se {r.err3 / 3 / r.value:.3e})", flush=True)
    assert rel(v, r.value) < 2e-5


@pytest.mark.parametrize("d,m,skip", [(150, 1500, None), (300, 1000, 3 * 2 ** 20 + 77),
                                      (120, 700, 2 ** 31 - 5000)])
def test_sobol_fp64_mfma(dbg, oracle, d, m, skip):
    """FP64 Sobol on the matrix-core kernel, whose walk takes its points from
    per-sub-block LDS tables (low Gray-code byte per lane, high part per tile):
    equal to the oracle's direct XOR of direction integers.  Tiles straddle
    256-point boundaries (m and the skips are not multiples of 256), and the
    large skips set high bits of the index."""
    P = W.dense_random(d, m=m, n_reps=2, seed=d, gen="Sobol")
    print(f"sobol fp64 mfma d={d} m={m} skip={skip}", flush=True)
    opts = dbg.QMC_opts(np.float64, m=P.m, n_reps=P.n_reps, max_pts=P.m, max_abs_err=0.0)
    D = dbg.QMCData(P.Sigma, P.a, P.b, opts, None, "Sobol", shifts=P.shifts, sobol_skip=skip)
    v, _, _ = dbg.qmc_pnorm(D)
    r = oracle_same_inputs(oracle, D)
    assert rel(v, r.value) < F64_TOL, (v, r.value)
    assert np.allclose(D.qmc_reps, r.rep_means, rtol=F64_TOL, atol=0)


@pytest.mark.parametrize("d,family,gen", [(150, "random", "Richtmyer"), (300, "random", "Richtmyer"),
                                          (700, "factor", "Richtmyer"), (1100, "factor", "Richtmyer"),
                                          (300, "random", "Sobol"), (700, "factor", "Sobol")])
def test_split16_contraction(dbg, oracle, d, family, gen):
    """The mixed dense kernel with its panel contraction on FP16 matrix cores
    (high/low FP16 split of both operands, three products, FP32 accumulation)
    against the FP64 oracle on the same factor, points and shifts, and against
    the FP32 matrix-core contraction.  The shapes cover a panel without
    contraction (d = 150), several panels with partial last panels and
    sub-blocks, and the deep lower orthant of the factor model whose
    estimate is far below the FP64 range (compared in log P).  No tile is
    re-evaluated in FP64 (at these point counts that would replace every
    tile)."""
    if family == "random":
        P = W.dense_random(d, m=2048, n_reps=3, seed=d, gen=gen)
    else:
        P = W.factor_model(d, m=2048, n_reps=2, seed=d, gen=gen)
    print(f"split16 d={d} {family} {gen}", flush=True)
    res = {}
    opts = dbg.QMC_opts(np.float32, m=P.m, n_reps=P.n_reps, max_pts=P.m, max_abs_err=0.0)
    for c in ("split16", "fp32"):
        D = dbg.QMCData(P.Sigma, P.a, P.b, opts, None, gen, shifts=P.shifts, contraction=c, refine_tiles=0)
        dbg.qmc_pnorm(D)
        res[c] = (D.last_result.log_value, np.array(D.qmc_reps, dtype=np.float64))
    D64 = _plan(dbg, P, gen=gen, phi_form="tail")
    r = oracle_same_inputs(oracle, D64)
    ls, lf, lo = res["split16"][0], res["fp32"][0], r.log_value
    print(f"  log P split16 {ls:.10f} fp32 {lf:.10f} oracle {lo:.10f}", flush=True)
    assert np.isfinite(ls)
    # |delta log P| is the relative error of P, bounded absolutely for both
    # contractions: 2e-5 where P is moderate (measured up to 3.2e-6), 1e-3 in
    # the deep lower orthant of the factor model (log P about -2e4 at d = 700
    # and -3e4 at d = 1100; measured up to 4.7e-4)
    tol = 2e-5 if family == "random" else 1e-3
    assert abs(ls - lo) < tol and abs(lf - lo) < tol, (ls, lf, lo)


@pytest.mark.parametrize("d,m", [(96, 700), (130, 300), (257, 1000), (400, 513)])
def test_split16_edge_shapes(dbg, oracle, d, m):
    """Split-FP16 kernel at the edges of its blocking: d = 96 (smallest matrix-
    core shape, one panel, no contraction), a second panel of one dimension
    (d = 130), the first panel with eight factor chunks (d = 257), and point
    counts that leave part of the last 256-sample tile empty.  No tile is
    re-evaluated in FP64."""
    P = W.dense_random(d, m=m, n_reps=2, seed=d + 1)
    print(f"split16 edge d={d} m={m}", flush=True)
    opts = dbg.QMC_opts(np.float32, m=P.m, n_reps=P.n_reps, max_pts=P.m, max_abs_err=0.0)
    vals = {}
    for c in ("split16", "fp32"):
        D = dbg.QMCData(P.Sigma, P.a, P.b, opts, None, "Richtmyer", shifts=P.shifts, contraction=c,
                        refine_tiles=0)
        vals[c], _, _ = dbg.qmc_pnorm(D)
    r = oracle_same_inputs(oracle, _plan(dbg, P, phi_form="tail"))
    print(f"  split16 {vals['split16']:.12g} fp32 {vals['fp32']:.12g} oracle {r.value:.12g}", flush=True)
    # measured 1.8e-7 to 2.8e-6
    assert rel(vals["split16"], r.value) < 2e-5 and rel(vals["fp32"], r.value) < 2e-5
