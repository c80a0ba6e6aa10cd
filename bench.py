#!/usr/bin/env python3
"""Benchmark of the Genz QMC multivariate-normal path on one MI355X.

Workload (the headline configuration): dense factor model, d = 2500,
Sigma = standardize(B B^T + diag(psi)) with B 2500 x 8, lower orthant
b = Phi^{-1}(u), u ~ U(0.01, 0.10); Richtmyer lattice, N = 2^18 points x 10
random shifts.  One step is one whole job with the inputs resident in HBM:
the device Genz-Bretz Cholesky of Sigma, truncate/pack, the fused
point-generation + integration kernel over all N x 10 samples and the FP64
shift reductions.

Reported: sample*dims/s of the FP32-mixed path (the value), the plain FP64
path on the same job, their ratio, the relative error of each against the
FP64 CPU oracle on a bounded sample of the same points and shifts, the
roofline of the integration kernel (device time from HIP events on the
plan's stream) and the CPU oracle's own throughput on that sample.

Both paths run the tail form of the normal CDF (complementary tail masses,
no [-9, 9] clamp): at this shape the reference's erf and erfc forms give every
sample weight exactly 0 (per-dimension masses fall below the clamp), so their
estimate is 0; the tail form keeps log P finite.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 matrix
PEAK_FP64_TFLOPS = 78.6
PEAK_FP16_TFLOPS = 2500.0  # MI355X dense FP16 matrix
PEAK_HBM_GBPS = 8000.0

# HBM bytes per launch of the mixed dense integration kernel, keyed by
# (contraction, d, N, shifts): rocprofv3 FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00
# (factors from a plain 1 GiB device copy), profiles/dense_d2500_mixed_counters.md.
# Other shapes report null.
TRAFFIC_BYTES = {("split16", 2500, 2 ** 18, 10): 2.860e11, ("fp32", 2500, 2 ** 18, 10): 2.839e11}


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--d", type=int, default=2500)
    ap.add_argument("--m", type=int, default=2 ** 18)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--cpu-points", type=int, default=2 ** 18,
                    help="points of shift 1 in the CPU oracle sample (2^18: the whole first shift, "
                         "several seconds on 16 host threads)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-grad", action="store_true", help="skip the gradient leg (d = 1000 factor model)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched leg (4096 problems)")
    ap.add_argument("--grad-only", action="store_true",
                    help="only run the gradient leg and print its record (for profiling it alone)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only run the batched leg and print its record (for profiling it alone)")
    ap.add_argument("--only", choices=("mixed", "fp64", "sparse", "chol"), default=None,
                    help="only run one leg at the headline shape (mixed or fp64 dense job, the banded d = 4000 "
                         "sparse job in both precisions, or the d = 2500 Cholesky alone) and print its timing, "
                         "for profiling it alone")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other configurations (d = 16, d = 500, Sobol at d = 2500, banded d = 4000)")
    ap.add_argument("--contraction", default="auto", choices=("auto", "fp32", "split16"),
                    help="arithmetic of the mixed dense panel contraction (auto: split16 where built)")
    ap.add_argument("--refine", type=int, default=-1,
                    help="tiles per shift re-evaluated in FP64 after the mixed round (-1: the library's default)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes three times (the reference "
                         "kernel for scaling the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the path would shard "
              "by (shift, point range) with one reduction of the shift sums)",
              file=sys.stderr)
        return 2
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)

    import numpy as np

    import dbgenzpaper_amd as dbg
    from dbgenzpaper_amd import workloads as W

    if args.grad_only:
        print(json.dumps({"adjoint": grad_leg(args.steps)}), flush=True)
        return 0
    if args.batch_only:
        print(json.dumps({"batched": batch_leg(args.steps)}), flush=True)
        return 0
    if args.only:
        print(json.dumps(single_leg(args.only, args.steps, args.warmup, args.d, args.m, args.reps)), flush=True)
        return 0

    d, m, R = args.d, args.m, args.reps
    _log(f"workload factor_model d={d} k=8 m={m} reps={R}")
    P = W.factor_model(d, 8, m=m, n_reps=R, seed=1)

    def make(T, m_, R_, shifts, contraction=None, refine=None):
        opts = dbg.QMC_opts(T, m=m_, n_reps=R_, max_pts=m_, max_abs_err=0.0)
        return dbg.QMCData(P.Sigma, P.a, P.b, opts, None, "Richtmyer", shifts=shifts, phi_form="tail",
                           contraction=contraction, refine_tiles=refine)

    def run(D, steps, warmup):
        for _ in range(warmup):
            D.refactor()
            D.loop(0)
            D.sync()
        kms, rms = [], []
        D.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            D.refactor()
            D.loop(0)
            kms.append(D.last_kernel_ms())
            rms.append(D.refine_info()[2])
        D.sync()
        t1 = time.perf_counter()
        return (t1 - t0) / steps, sum(kms) / len(kms), D.summary(), sum(rms) / len(rms), D.refine_info()

    contraction = "fp32" if args.contraction == "fp32" else "split16"  # auto: split16 at this shape
    Dm = make(np.float32, m, R, P.shifts, args.contraction, args.refine)
    _log("mixed plan ready")
    t_mixed, k_mixed, s_mixed, r_ms, (r_k, r_share, _) = run(Dm, args.steps, args.warmup)
    _log(f"mixed: {t_mixed * 1e3:.1f} ms/step, kernel {k_mixed:.1f} ms, FP64 re-evaluation {r_ms:.1f} ms "
         f"({r_k} tiles/shift, share left {r_share:.1e}), log P {s_mixed.log_value:.6f}")
    sdims = float(m) * R * d
    value = sdims / t_mixed
    out = {
        "metric": "sample·dims/s at d=2500 mixed vs plain FP64 (rel err≤1e-5); batched probs/s",
        "value": value,
        "unit": "sample*dims/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("fp32 (contraction as high/low FP16 products with FP32 accumulation, special functions, "
                  "weights) + fp64 (factor, points of the re-evaluated tiles, their walk, reductions)"
                  if contraction == "split16" else
                  "fp32 (contraction, special functions) + fp64 (factor, re-evaluated tiles, reductions)"),
        "data": "synthetic (seeded factor model, random-init inputs of the configured shape)",
        "config": {"workload": f"factor_model d={d} k=8 lower orthant u~U(0.01,0.10), Richtmyer "
                               f"N={m} x {R} shifts, step = Cholesky + mixed integration + FP64 re-evaluation "
                               f"of the {r_k} heaviest tiles per shift + reductions",
                   "d": d, "points": m, "shifts": R, "parallelism": "single GPU"},
        "log_p_mixed": s_mixed.log_value,
        "kernel_ms_mixed": k_mixed,
        "refine": {"tiles_per_shift": r_k, "ms": r_ms, "share_left_in_fp32": r_share},
        "contraction": contraction,
    }
    flops = float(m) * R * d * (d - 1)
    achieved = flops / (k_mixed * 1e-3) / 1e12
    traffic = TRAFFIC_BYTES.get((contraction, d, m, R))
    # roofline of the dominant kernel (the mixed integration): the algorithmic
    # d(d-1) flops per sample over its HIP-event time, against the peak of the
    # pipe it runs on -- FP16 matrix cores for the split contraction (which
    # executes each multiply-add as three FP16 products), FP32 matrix cores
    # for the fp32 contraction; its HBM rate against 8 TB/s beside it
    peak = PEAK_FP16_TFLOPS if contraction == "split16" else PEAK_FP32_TFLOPS
    out["roofline"] = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                       "frac": achieved / peak, "traffic": traffic,
                       "kernel": f"dense integration (mixed, {contraction} contraction)",
                       "flops_per_launch": flops,
                       "executed_fp16_frac": (3 * achieved / PEAK_FP16_TFLOPS) if contraction == "split16" else None,
                       "fp32_peak_ratio": achieved / PEAK_FP32_TFLOPS}
    if traffic:
        out["roofline"]["hbm"] = {"achieved_GBps": traffic / (k_mixed * 1e-3) / 1e9, "peak_GBps": PEAK_HBM_GBPS,
                                  "frac": traffic / (k_mixed * 1e-3) / 1e9 / PEAK_HBM_GBPS}
    Dm.close()

    if not args.no_fp64:
        # the same job without the FP64 re-evaluation (the FP32 walk's own error)
        Du = make(np.float32, m, R, P.shifts, args.contraction, 0)
        t_u, k_u, s_u, _, _ = run(Du, max(1, args.steps), min(args.warmup, 1))
        Du.close()
        _log(f"mixed, no FP64 re-evaluation: {t_u * 1e3:.1f} ms/step, log P {s_u.log_value:.6f}")
        out["mixed_unrefined"] = {"ms_per_step": t_u * 1e3, "kernel_ms": k_u, "log_p": s_u.log_value}
        if contraction == "split16":
            # the same job with the FP32 matrix-core contraction, for comparison
            Df = make(np.float32, m, R, P.shifts, "fp32", args.refine)
            t_f, k_f, s_f, _, _ = run(Df, max(1, args.steps), min(args.warmup, 1))
            _log(f"mixed (fp32 contraction): {t_f * 1e3:.1f} ms/step, kernel {k_f:.1f} ms, log P {s_f.log_value:.6f}")
            out["fp32_contraction"] = {"ms_per_step": t_f * 1e3, "kernel_ms": k_f, "log_p": s_f.log_value}
            Df.close()
        D64 = make(np.float64, m, R, P.shifts)
        _log("fp64 plan ready")
        t64, k64, s64, _, _ = run(D64, max(1, args.steps), min(args.warmup, 1))
        _log(f"fp64: {t64 * 1e3:.1f} ms/step, kernel {k64:.1f} ms, log P {s64.log_value:.6f}")
        out["fp64_value"] = sdims / t64
        out["fp64_ms_per_step"] = t64 * 1e3
        out["kernel_ms_fp64"] = k64
        out["log_p_fp64"] = s64.log_value
        out["roofline_fp64"] = {"bound": "mfma", "achieved": flops / (k64 * 1e-3) / 1e12, "peak": PEAK_FP64_TFLOPS,
                                "unit": "TFLOP/s", "frac": flops / (k64 * 1e-3) / 1e12 / PEAK_FP64_TFLOPS}
        out["speedup_mixed_vs_fp64"] = t64 / t_mixed
        out["rel_err_mixed_vs_fp64_full"] = abs(math.expm1(s_mixed.log_value - s64.log_value))
        out["mixed_unrefined"]["rel_err_vs_fp64_full"] = abs(math.expm1(s_u.log_value - s64.log_value))
        out["mixed_unrefined"]["speedup_vs_fp64"] = t64 / t_u
        if "fp32_contraction" in out:
            fc = out["fp32_contraction"]
            fc["rel_err_vs_fp64_full"] = abs(math.expm1(fc["log_p"] - s64.log_value))
            fc["speedup_vs_fp64"] = t64 * 1e3 / fc["ms_per_step"]
        D64.close()

    if not args.no_batch:
        out["batched"] = batch_leg(args.steps)

    if not args.no_grad:
        out["adjoint"] = grad_leg(args.steps)

    if not args.no_configs:
        out["other_configs"] = config_legs(args.steps)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_leg(P, d, args.cpu_points, out, args.refine)

    print(json.dumps(out), flush=True)
    return 0


def single_leg(which, steps, warmup, d, m, R):
    """One leg alone: the mixed or FP64 dense job (factorization, integration,
    reductions) of the headline shape, the banded d = 4000 sparse job (mixed
    then FP64), or the headline shape's device Cholesky (refactor only)."""
    import numpy as np

    import dbgenzpaper_amd as dbg
    from dbgenzpaper_amd import workloads as W

    if which == "sparse":
        P = W.banded(4000, 16, m=m, n_reps=R, seed=2)
        out = {}
        for name, T, form in (("mixed", np.float32, "tail"), ("fp64", np.float64, "erf")):
            opts = dbg.QMC_opts(T, m=P.m, n_reps=P.n_reps, max_pts=P.m, max_abs_err=0.0)
            D = dbg.QMCDataSparse(P.Sigma, P.a, P.b, opts, None, "Richtmyer", shifts=P.shifts, phi_form=form)
            step, kms, sm = _time_loop(D, steps)
            out[name] = {"step_ms": step, "kernel_ms": kms, "log_p": sm.log_value}
            D.close()
        return {"leg": "banded d=4000 bw=16", **out}
    P = W.factor_model(d, 8, m=m, n_reps=R, seed=1)
    T = np.float64 if which in ("fp64", "chol") else np.float32
    opts = dbg.QMC_opts(T, m=m, n_reps=R, max_pts=m, max_abs_err=0.0)
    D = dbg.QMCData(P.Sigma, P.a, P.b, opts, None, "Richtmyer", shifts=P.shifts, phi_form="tail")
    if which == "chol":
        for _ in range(warmup):
            D.refactor()
        D.sync()
        t0 = time.perf_counter()
        for _ in range(max(1, steps)):
            D.refactor()
        D.sync()
        res = {"leg": f"cholesky_genz! d={d} (refactor)", "ms": (time.perf_counter() - t0) / max(1, steps) * 1e3}
    else:
        step, kms, sm = _time_loop(D, steps)
        res = {"leg": f"{which} dense d={d}", "step_ms": step, "kernel_ms": kms, "log_p": sm.log_value}
    D.close()
    return res


def batch_leg(steps):
    """Batched uneven sizes: 4096 independent problems, d_k ~ U{2..256},
    random correlation matrices (G G^T rescaled), a = -inf, b ~ U(-0.5, 2.5),
    Richtmyer N = 2^14 x 4 shifts each, in random arrival order; one
    graph-captured job (device factorizations, one persistent integration
    launch, per-problem reductions), FP32-mixed.  Worst relative error
    against the same batch solved in FP64, which is itself checked against
    single-problem FP64 plans on 16 of the problems."""
    import numpy as np

    import dbgenzpaper_amd as dbg
    from dbgenzpaper_amd import workloads as W
    from dbgenzpaper_amd.batch import Batch

    K, m, R = 4096, 2 ** 14, 4
    probs = W.batch_uneven(K, d_min=2, d_max=256, m=m, n_reps=R, seed=3)
    args = ([p.Sigma for p in probs], [p.a for p in probs], [p.b for p in probs], [p.shifts for p in probs])
    sdims = float(sum(p.d for p in probs)) * m * R

    def timed(T):
        t0 = time.perf_counter()
        B = Batch(*args, T=T, m=m, n_reps=R)
        t_create = time.perf_counter() - t0
        r = B.run()  # warm-up replay
        walls, dev = [], []
        for _ in range(max(1, steps)):
            t0 = time.perf_counter()
            r = B.run()
            walls.append(time.perf_counter() - t0)
            dev.append(B.last_ms())
        tune = B.tuning()
        B.close()
        return r, sum(walls) / len(walls), sum(dev) / len(dev), t_create, tune

    r32, w32, d32, c32, tune32 = timed(np.float32)
    _log(f"batched mixed: {w32 * 1e3:.1f} ms per job ({K / w32:.0f} probs/s), device {d32:.1f} ms, "
         f"width {tune32[0]}")
    r64, w64, d64, c64, tune64 = timed(np.float64)
    _log(f"batched fp64: {w64 * 1e3:.1f} ms per job, device {d64:.1f} ms")
    fin = np.isfinite(r32.log_value) & np.isfinite(r64.log_value)
    err = np.abs(np.expm1(r32.log_value[fin] - r64.log_value[fin]))
    nonfinite = [{"index": int(i), "d": int(probs[i].d), "log_p_mixed": float(r32.log_value[i]),
                  "log_p_fp64": float(r64.log_value[i]), "value_fp64": float(r64.value[i]),
                  "rank_mixed": int(r32.rank[i]), "rank_fp64": int(r64.rank[i])}
                 for i in np.flatnonzero(~fin)[:8]]
    if nonfinite:
        _log(f"batched: {int((~fin).sum())} problems with a non-finite log P: {nonfinite}")
    # the FP64 batch against single-problem FP64 plans (the per-problem solve of qmc_pnorm!)
    spot = []
    for idx in np.linspace(0, K - 1, 16).astype(int):
        p = probs[idx]
        opts = dbg.QMC_opts(np.float64, m=m, n_reps=R, max_pts=m, max_abs_err=0.0)
        D = dbg.QMCData(p.Sigma, p.a, p.b, opts, None, "Richtmyer", shifts=p.shifts, phi_form="erf")
        v, _, _ = dbg.qmc_pnorm(D)
        D.close()
        spot.append(abs(r64.value[idx] / v - 1.0))
    return {"workload": f"{K} problems, d~U{{2..256}} (sum d = {sum(p.d for p in probs)}), random correlation, "
                        f"a=-inf, b~U(-0.5,2.5), Richtmyer N={m} x {R} shifts, random arrival order, FP32-mixed, "
                        "one graph (factorizations + persistent integration + reductions)",
            "probs_per_s": K / w32, "ms_per_job": w32 * 1e3, "device_ms": d32,
            "sample_dims_per_s": sdims / w32,
            "worst_rel_err_vs_fp64": float(err.max()), "median_rel_err_vs_fp64": float(np.median(err)),
            "nonfinite_count": int((~fin).sum()), "nonfinite_first": nonfinite,
            "fp64_probs_per_s": K / w64, "fp64_ms_per_job": w64 * 1e3,
            "fp64_batch_vs_single_fp64_plans_max_rel": float(max(spot)),
            "autotune_width": tune32[0], "autotune_ms_w8_w16_w32": tune32[1],
            "create_s_mixed": c32, "create_s_fp64": c64}


def _time_loop(D, steps):
    """Device time of the integration kernel (HIP events) and wall time of a
    whole step (refactor + loop), averaged over steps after one warm-up."""
    D.refactor()
    D.loop(0)
    D.sync()
    ks, t0 = [], time.perf_counter()
    for _ in range(max(1, steps)):
        D.refactor()
        D.loop(0)
        ks.append(D.last_kernel_ms())
    D.sync()
    return (time.perf_counter() - t0) / max(1, steps) * 1e3, sum(ks) / len(ks), D.summary()


def config_legs(steps):
    """The other configurations of the benchmark set, each on its own shape:
    d = 16 random correlation (N = 2^14 x 8, FP64, with the CPU oracle on the
    whole job), the factor model at d = 500 (mixed and FP64), the Sobol
    variant at d = 2500 (mixed and FP64), and the banded d = 4000 sparse path
    (bandwidth 16, unpivoted, N = 2^18 x 10; FP64 erf form and mixed)."""
    import numpy as np

    import dbgenzpaper_amd as dbg
    from dbgenzpaper_amd import workloads as W

    res = {}

    def dense_pair(P, gen, label):
        out = {}
        for name, T in (("mixed", np.float32), ("fp64", np.float64)):
            opts = dbg.QMC_opts(T, m=P.m, n_reps=P.n_reps, max_pts=P.m, max_abs_err=0.0)
            D = dbg.QMCData(P.Sigma, P.a, P.b, opts, None, gen, shifts=P.shifts, phi_form="tail")
            step, kms, sm = _time_loop(D, steps)
            out[name] = {"step_ms": step, "kernel_ms": kms, "log_p": sm.log_value,
                         "sample_dims_per_s": P.m * P.n_reps * P.Sigma.shape[0] / (step * 1e-3)}
            D.close()
        out["rel_err_mixed_vs_fp64"] = abs(math.expm1(out["mixed"]["log_p"] - out["fp64"]["log_p"]))
        _log(f"{label}: mixed kernel {out['mixed']['kernel_ms']:.2f} ms, fp64 {out['fp64']['kernel_ms']:.2f} ms, "
             f"rel err {out['rel_err_mixed_vs_fp64']:.2e}")
        return out

    # d = 16 random correlation, b ~ U(0, 2), N = 2^14 x 8 (the CPU-runnable case)
    P0 = W.dense_random(16)
    opts = dbg.QMC_opts(np.float64, m=P0.m, n_reps=P0.n_reps, max_pts=P0.m, max_abs_err=0.0)
    D0 = dbg.QMCData(P0.Sigma, P0.a, P0.b, opts, None, "Richtmyer", shifts=P0.shifts)
    t0 = time.perf_counter()
    for _ in range(max(1, steps)):
        v0, e0, _ = dbg.qmc_pnorm(D0)
    t_gpu = (time.perf_counter() - t0) / max(1, steps)
    from oracle import oracle as O

    def oracle_job():
        # the whole job on the CPU: the constructor's factorization, then
        # qmc_pnorm! with the clamped block sizes (src/qmc.jl:166-174)
        F0 = O.prepare_dense(P0.Sigma, P0.a, P0.b)
        bs_i, bs_i2 = O.clamp_block_sizes(16, 64, 32)
        return O.qmc_pnorm(F0.U, F0.a, F0.b, m=P0.m, n_reps=P0.n_reps, shifts=P0.shifts, bs_i=bs_i, bs_i2=bs_i2)

    oracle_job()  # loads the library and starts its threads
    t0 = time.perf_counter()
    r0 = oracle_job()
    t_cpu = time.perf_counter() - t0
    res["d16_random_correlation"] = {"value": v0, "err3": e0, "gpu_ms": t_gpu * 1e3, "cpu_oracle_ms": t_cpu * 1e3,
                                     "cpu_threads": O.num_threads(), "rel_err_vs_oracle": abs(v0 / r0.value - 1)}
    D0.close()
    _log(f"d=16: P {v0:.6f} +- {e0:.1e}, gpu {t_gpu * 1e3:.2f} ms/call, oracle {t_cpu * 1e3:.1f} ms")

    res["factor_model_d500"] = dense_pair(W.factor_model(500, 8, m=2 ** 18, n_reps=10, seed=1), "Richtmyer", "d=500")
    res["factor_model_d2500_sobol"] = dense_pair(W.factor_model(2500, 8, m=2 ** 18, n_reps=10, seed=1, gen="Sobol"),
                                                 "Sobol", "d=2500 Sobol")

    # banded d = 4000, bandwidth 16, unpivoted sparse path
    P2 = W.banded(4000, 16, m=2 ** 18, n_reps=10, seed=2)
    out = {}
    for name, T, form in (("mixed", np.float32, "tail"), ("fp64", np.float64, "erf")):
        opts = dbg.QMC_opts(T, m=P2.m, n_reps=P2.n_reps, max_pts=P2.m, max_abs_err=0.0)
        D = dbg.QMCDataSparse(P2.Sigma, P2.a, P2.b, opts, None, "Richtmyer", shifts=P2.shifts, phi_form=form)
        step, kms, sm = _time_loop(D, steps)
        out[name] = {"step_ms": step, "kernel_ms": kms, "log_p": sm.log_value,
                     "sample_dims_per_s": P2.m * P2.n_reps * 4000 / (step * 1e-3)}
        D.close()
    out["rel_err_mixed_vs_fp64"] = abs(math.expm1(out["mixed"]["log_p"] - out["fp64"]["log_p"]))
    _log(f"banded d=4000: mixed kernel {out['mixed']['kernel_ms']:.2f} ms, fp64 {out['fp64']['kernel_ms']:.2f} ms, "
         f"rel err {out['rel_err_mixed_vs_fp64']:.2e}")
    res["banded_d4000_bw16"] = out
    return res


def grad_leg(steps):
    """Gradient of log P on its own workload: factor model d = 1000,
    B 1000 x 16, lower orthant b = Phi^{-1}(u), u ~ U(0.02, 0.20), Richtmyer
    N = 2^18 x 8, FP32-mixed.  Device time of one forward integration kernel
    against the device time of the whole gradient (forward with records,
    reverse sweep, Y^T Sbar, Cholesky / standardization chain to Sigma)."""
    import numpy as np

    import dbgenzpaper_amd as dbg
    from dbgenzpaper_amd import workloads as W

    d, m, R = 1000, 2 ** 18, 8
    P = W.factor_model(d, 16, m=m, n_reps=R, seed=4, u_range=(0.02, 0.20))
    opts = dbg.QMC_opts(np.float32, m=m, n_reps=R, max_pts=m, max_abs_err=0.0)
    D = dbg.QMCData(P.Sigma, P.a, P.b, opts, None, "Richtmyer", shifts=P.shifts, phi_form="tail")
    D.loop(0)
    D.sync()
    fw = []
    for _ in range(max(1, steps)):
        D.loop(0)
        D.sync()
        fw.append(D.last_kernel_ms())
    g = D.grad()  # first call allocates the records
    gm = []
    for _ in range(max(1, steps)):
        g = D.grad()
        gm.append(D.last_kernel_ms())
    fwd = sum(fw) / len(fw)
    grd = sum(gm) / len(gm)
    chunk, n_chunks, dev_bytes = D.grad_info()
    _log(f"adjoint d={d}: forward kernel {fwd:.1f} ms, gradient {grd:.1f} ms, log P {g.log_value:.4f}, "
         f"{n_chunks} chunks of {chunk} samples, {dev_bytes / 2**30:.2f} GiB")
    D.close()
    # the forward with the FP32 matrix-core contraction (the arithmetic of the
    # gradient's own record and reverse sweeps)
    Df = dbg.QMCData(P.Sigma, P.a, P.b, opts, None, "Richtmyer", shifts=P.shifts, phi_form="tail",
                     contraction="fp32")
    Df.loop(0)
    Df.sync()
    ff = []
    for _ in range(max(1, steps)):
        Df.loop(0)
        Df.sync()
        ff.append(Df.last_kernel_ms())
    Df.close()
    fwd32 = sum(ff) / len(ff)
    # the gradient re-runs the forward (with records) first: the adjoint proper
    # is the rest, estimated as the total minus one forward kernel
    return {"workload": f"factor_model d={d} k=16 lower orthant u~U(0.02,0.20), Richtmyer N={m} x {R} shifts, "
                        "FP32-mixed, gradient w.r.t. a, b, Sigma",
            "forward_kernel_ms": fwd, "gradient_ms": grd, "ratio": grd / fwd,
            "adjoint_only_ms": grd - fwd, "adjoint_only_ratio": (grd - fwd) / fwd, "log_p": g.log_value,
            "forward_kernel_ms_fp32_contraction": fwd32, "ratio_vs_fp32_contraction_forward": grd / fwd32,
            "record_chunk_samples": chunk, "record_chunks": n_chunks, "device_bytes": dev_bytes,
            "records_unchunked_bytes": 2 * (d - 1) * m * R * 4,
            "grad_b_norm": float(np.linalg.norm(g.b)), "grad_sigma_norm": float(np.linalg.norm(g.Sigma))}


def copy_calibration(nbytes):
    """A plain device-to-device copy of nbytes by the HIP runtime (its blit
    kernel), run three times: under rocprofv3 --pmc the summed FETCH_SIZE and
    WRITE_SIZE of its dispatches give the counters' scale against known
    bytes."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    x, y = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(x), nbytes) == 0 and hip.hipMalloc(ctypes.byref(y), nbytes) == 0
    assert hip.hipMemset(x, 1, nbytes) == 0
    for _ in range(3):
        assert hip.hipMemcpy(y, x, nbytes, 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(x)
    hip.hipFree(y)
    print(json.dumps({"copy_bytes": nbytes, "copies": 3}), flush=True)
    return 0


def cpu_leg(P, d, n_pts, out, refine):
    """The CPU baseline and the end-to-end check against the oracle.

    The FP64 oracle (C + OpenMP over all host threads; the QMCData
    constructor's cholesky_genz! on its own, then qmc_loop! in the tail form
    with the reference's blocking bs_j = 2^7, bs_i = 2^10, bs_i2 = 2^6) runs
    the first n_pts points of shift 1.  Its factorization is its own, so the
    device estimates of the same points (FP64 and mixed plans, each factoring
    on the device) are compared end to end.  The CPU rate of the whole job is
    the measured Cholesky plus R times the measured shift."""
    import numpy as np

    import dbgenzpaper_amd as dbg
    from oracle import oracle as O

    sh = np.asfortranarray(P.shifts[:, :1])
    R = P.n_reps
    threads = O.num_threads()
    _log(f"cpu oracle: Cholesky + {n_pts} points of one shift, {threads} threads")
    O.qmc_pnorm(np.eye(4), np.full(4, -1.0), np.ones(4), m=64, n_reps=1, shifts=np.zeros((3, 1)))  # load, threads up
    t0 = time.perf_counter()
    F = O.prepare_dense(P.Sigma, P.a, P.b)
    t_chol = time.perf_counter() - t0
    t0 = time.perf_counter()
    r = O.qmc_pnorm(F.U, F.a, F.b, m=n_pts, n_reps=1, shifts=sh, phi_form="tail", bs_j=2 ** 7, bs_i=2 ** 10,
                    bs_i2=2 ** 6)
    t_shift = time.perf_counter() - t0
    o64 = dbg.QMC_opts(np.float64, m=n_pts, n_reps=1, max_pts=n_pts, max_abs_err=0.0)
    D64 = dbg.QMCData(P.Sigma, P.a, P.b, o64, None, "Richtmyer", shifts=sh, phi_form="tail")
    o32 = dbg.QMC_opts(np.float32, m=n_pts, n_reps=1, max_pts=n_pts, max_abs_err=0.0)
    Dm = dbg.QMCData(P.Sigma, P.a, P.b, o32, None, "Richtmyer", shifts=sh, refine_tiles=refine)
    dbg.qmc_pnorm(D64)
    dbg.qmc_pnorm(Dm)
    out["rel_err_fp64_vs_oracle"] = abs(math.expm1(D64.last_result.log_value - r.log_value))
    out["rel_err_mixed_vs_oracle"] = abs(math.expm1(Dm.last_result.log_value - r.log_value))
    out["oracle_sample_log_p"] = r.log_value
    out["oracle_sample"] = f"first {n_pts} points of shift 1, oracle's own factorization (end to end)"
    D64.close()
    Dm.close()
    job_s = t_chol + R * t_shift * (P.m / n_pts)
    return {"value": float(P.m) * R * d / job_s, "unit": "sample*dims/s", "cores": threads, "kind": "port",
            "sample": f"cholesky_genz! of the d={d} problem ({t_chol:.2f} s) and the first {n_pts} Richtmyer points "
                      f"of shift 1 ({t_shift:.2f} s), FP64 tail form; value = whole job ({R} shifts of {P.m} "
                      f"points) at those rates, {job_s:.1f} s",
            "cholesky_s": t_chol, "shift_s": t_shift * (P.m / n_pts),
            "loop_sample_dims_per_s": n_pts * d / t_shift}


if __name__ == "__main__":
    sys.exit(main())
