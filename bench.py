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

# HBM bytes per launch of the mixed dense integration kernel, keyed by
# (contraction, d, N, shifts): rocprofv3 FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00
# (factors from a plain 1 GiB device copy), profiles/dense_d2500_mixed_counters.md.
# Other shapes report null.
TRAFFIC_BYTES = {("split16", 2500, 2 ** 18, 10): 2.864e11, ("fp32", 2500, 2 ** 18, 10): 2.839e11}


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
    ap.add_argument("--grad-only", action="store_true",
                    help="only run the gradient leg and print its record (for profiling it alone)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other configurations (d = 16, d = 500, Sobol at d = 2500, banded d = 4000)")
    ap.add_argument("--contraction", default="auto", choices=("auto", "fp32", "split16"),
                    help="arithmetic of the mixed dense panel contraction (auto: split16 where built)")
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

    d, m, R = args.d, args.m, args.reps
    _log(f"workload factor_model d={d} k=8 m={m} reps={R}")
    P = W.factor_model(d, 8, m=m, n_reps=R, seed=1)

    def make(T, m_, R_, shifts, contraction=None):
        opts = dbg.QMC_opts(T, m=m_, n_reps=R_, max_pts=m_, max_abs_err=0.0)
        return dbg.QMCData(P.Sigma, P.a, P.b, opts, None, "Richtmyer", shifts=shifts, phi_form="tail",
                           contraction=contraction)

    def run(D, steps, warmup):
        for _ in range(warmup):
            D.refactor()
            D.loop(0)
            D.sync()
        kms = []
        D.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            D.refactor()
            D.loop(0)
            kms.append(D.last_kernel_ms())
        D.sync()
        t1 = time.perf_counter()
        return (t1 - t0) / steps, sum(kms) / len(kms), D.summary()

    contraction = "fp32" if args.contraction == "fp32" else "split16"  # auto: split16 at this shape
    Dm = make(np.float32, m, R, P.shifts, args.contraction)
    _log("mixed plan ready")
    t_mixed, k_mixed, s_mixed = run(Dm, args.steps, args.warmup)
    _log(f"mixed: {t_mixed * 1e3:.1f} ms/step, kernel {k_mixed:.1f} ms, log P {s_mixed.log_value:.6f}")
    sdims = float(m) * R * d
    value = sdims / t_mixed
    out = {
        "metric": "sample*dims/s, FP32-mixed Genz QMC MVN probability at d=2500 (N=2^18 x 10 shifts)",
        "value": value,
        "unit": "sample*dims/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("fp32-equivalent contraction (FP16 matrix cores, high/low split of both operands, three "
                  "products, FP32 accumulation) + fp32 special functions + fp64 (factor, points, reductions)"
                  if contraction == "split16" else
                  "fp32 (contraction, special functions) + fp64 (factor, points, reductions)"),
        "data": "synthetic (seeded factor model, random-init inputs of the configured shape)",
        "config": {"workload": f"factor_model d={d} k=8 lower orthant u~U(0.01,0.10), Richtmyer "
                               f"N={m} x {R} shifts, step = Cholesky + integration + reductions",
                   "d": d, "points": m, "shifts": R, "parallelism": "single GPU"},
        "log_p_mixed": s_mixed.log_value,
        "kernel_ms_mixed": k_mixed,
        "contraction": contraction,
    }
    flops = float(m) * R * d * (d - 1)
    achieved = flops / (k_mixed * 1e-3) / 1e12
    # achieved: the algorithmic d(d-1) flops per sample (FP32-accurate multiply-adds)
    # against the FP32 matrix peak, the precision the contraction delivers; the
    # split form runs them as three FP16 products each on the FP16 pipe
    out["roofline"] = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                       "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                       "traffic": TRAFFIC_BYTES.get((contraction, d, m, R)),
                       "kernel": f"dense integration (mixed, {contraction} contraction)",
                       "flops_per_launch": flops}
    if contraction == "split16":
        out["roofline"]["fp16_pipe"] = {"executed_tflops": 3 * achieved, "peak": PEAK_FP16_TFLOPS,
                                        "frac": 3 * achieved / PEAK_FP16_TFLOPS}
    Dm.close()
    if contraction == "split16" and not args.no_fp64:
        # the same job with the FP32 matrix-core contraction, for comparison
        Df = make(np.float32, m, R, P.shifts, "fp32")
        t_f, k_f, s_f = run(Df, max(1, args.steps), min(args.warmup, 1))
        _log(f"mixed (fp32 contraction): {t_f * 1e3:.1f} ms/step, kernel {k_f:.1f} ms, log P {s_f.log_value:.6f}")
        out["fp32_contraction"] = {"ms_per_step": t_f * 1e3, "kernel_ms": k_f, "log_p": s_f.log_value}
        Df.close()

    if not args.no_fp64:
        D64 = make(np.float64, m, R, P.shifts)
        _log("fp64 plan ready")
        t64, k64, s64 = run(D64, max(1, args.steps), min(args.warmup, 1))
        _log(f"fp64: {t64 * 1e3:.1f} ms/step, kernel {k64:.1f} ms, log P {s64.log_value:.6f}")
        out["fp64_value"] = sdims / t64
        out["fp64_ms_per_step"] = t64 * 1e3
        out["kernel_ms_fp64"] = k64
        out["log_p_fp64"] = s64.log_value
        out["speedup_mixed_vs_fp64"] = t64 / t_mixed
        out["rel_err_mixed_vs_fp64_full"] = abs(math.expm1(s_mixed.log_value - s64.log_value))
        if "fp32_contraction" in out:
            fc = out["fp32_contraction"]
            fc["rel_err_vs_fp64_full"] = abs(math.expm1(fc["log_p"] - s64.log_value))
            fc["speedup_vs_fp64"] = t64 * 1e3 / fc["ms_per_step"]
        D64.close()

    if not args.no_grad:
        out["adjoint"] = grad_leg(args.steps)

    if not args.no_configs:
        out["other_configs"] = config_legs(args.steps)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_leg(P, d, args.cpu_points, out)

    print(json.dumps(out), flush=True)
    return 0


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
    from tests._helpers import oracle_same_inputs
    oracle_same_inputs(O, D0)  # loads the library and starts its threads
    t0 = time.perf_counter()
    r0 = oracle_same_inputs(O, D0)
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
    _log(f"adjoint d={d}: forward kernel {fwd:.1f} ms, gradient {grd:.1f} ms, log P {g.log_value:.4f}")
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
            "grad_b_norm": float(np.linalg.norm(g.b)), "grad_sigma_norm": float(np.linalg.norm(g.Sigma))}


def copy_calibration(nbytes):
    """A plain device-to-device copy of nbytes (torch hands it to the HIP
    runtime's blit kernel), run three times: under rocprofv3 --pmc the summed
    FETCH_SIZE and WRITE_SIZE of its dispatches give the counters' scale
    against known bytes."""
    import torch

    n = nbytes // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "check": float(y[-1].item())}), flush=True)
    return 0


def cpu_leg(P, d, n_pts, out):
    """FP64 oracle (tail form, the reference's blocking bs_j=2^7, bs_i=2^10,
    bs_i2=2^6) on the first n_pts points of shift 1, all host cores; the GPU
    paths run the same sample for their error against it."""
    import numpy as np

    import dbgenzpaper_amd as dbg
    from oracle import oracle as O
    from tests._helpers import oracle_same_inputs

    sh = np.asfortranarray(P.shifts[:, :1])
    o64 = dbg.QMC_opts(np.float64, m=n_pts, n_reps=1, max_pts=n_pts, max_abs_err=0.0,
                       block_size_i=2 ** 10, block_size_i2=2 ** 6, block_size_j=2 ** 7)
    D64 = dbg.QMCData(P.Sigma, P.a, P.b, o64, None, "Richtmyer", shifts=sh, phi_form="tail")
    o32 = dbg.QMC_opts(np.float32, m=n_pts, n_reps=1, max_pts=n_pts, max_abs_err=0.0)
    Dm = dbg.QMCData(P.Sigma, P.a, P.b, o32, None, "Richtmyer", shifts=sh)
    dbg.qmc_pnorm(D64)
    dbg.qmc_pnorm(Dm)
    threads = O.num_threads()
    _log(f"cpu oracle: {n_pts} points, {threads} threads")
    t0 = time.perf_counter()
    r = oracle_same_inputs(O, D64)
    dt = time.perf_counter() - t0
    lg64 = D64.last_result.log_value
    lgm = Dm.last_result.log_value
    out["rel_err_fp64_vs_oracle"] = abs(math.expm1(lg64 - r.log_value))
    out["rel_err_mixed_vs_oracle"] = abs(math.expm1(lgm - r.log_value))
    out["oracle_sample_log_p"] = r.log_value
    D64.close()
    Dm.close()
    return {"value": n_pts * d / dt, "unit": "sample*dims/s", "cores": threads, "kind": "port",
            "sample": f"first {n_pts} Richtmyer points of shift 1 of the same d={d} problem, FP64 tail form, "
                      f"{dt:.2f} s wall", "seconds": dt}


if __name__ == "__main__":
    sys.exit(main())
