"""Error budget of the mixed dense walk, by FP64 substitution (CPU study).

mixed_budget.cpp restates the mixed walk of the split-FP16 dense kernel for
one shift on the CPU -- 32-bit fixed-point points, the split-FP16 panel
contraction with FP32 accumulation, the FP32 in-panel update and in-block
dot, z = (b - s) inv in FP32, the FP32 tail mass and quantile
(dbgenzpaper_amd/csrc/tailq.h, host build) and the FP32 weight product -- with
each stage switchable to FP64.  Every stage run alone in FP32 against the
all-FP64 walk gives the stage's share of the error in log P; with all stages
in FP64 the emulator equals the FP64 oracle's tail form on the same points.
It also reports how concentrated the per-sample weights are and what
re-evaluating the heaviest samples in FP64 leaves.

    python tests/numerics/mixed_budget.py [d] [points]     (default 2500 2048)

Builds mixed_budget.so beside this file with g++ (OpenMP).
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

FIELDS = ["pt32", "pbits", "pacc", "pgroup", "inpanel", "inblock", "z", "mass", "quant", "weight", "state", "ybits"]
DEVICE = dict(pt32=1, pbits=22, pacc=1, pgroup=4, inpanel=1, inblock=1, z=1, mass=1, quant=1, weight=1, state=1,
              ybits=0)
FP64 = dict(pt32=0, pbits=0, pacc=0, pgroup=4, inpanel=0, inblock=0, z=0, mass=0, quant=0, weight=0, state=0, ybits=0)


class Cfg(C.Structure):
    _fields_ = [(f, C.c_int) for f in FIELDS]


def _lib():
    so = os.path.join(HERE, "mixed_budget.so")
    src = os.path.join(HERE, "mixed_budget.cpp")
    if not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
        subprocess.check_call(["g++", "-O2", "-fopenmp", "-std=c++17", "-shared", "-fPIC", "-o", so, src])
    return C.CDLL(so)


def main():
    from dbgenzpaper_amd import workloads as W
    from oracle import oracle as O

    d = int(sys.argv[1]) if len(sys.argv) > 1 else 2500
    m = int(sys.argv[2]) if len(sys.argv) > 2 else 2048
    lib = _lib()
    P = W.factor_model(d, 8, m=m, n_reps=1, seed=1)
    F = O.prepare_dense(P.Sigma, P.a, P.b)
    U = np.asfortranarray(np.triu(F.U))
    a, b = np.ascontiguousarray(F.a), np.ascontiguousarray(F.b)
    q = O.sqrt_primes(d - 1)
    sh = np.ascontiguousarray(P.shifts[:, 0])
    p = lambda x: x.ctypes.data_as(C.c_void_p)

    def run(cfg):
        lw = np.zeros(m)
        lib.budget_walk(C.c_int(d), p(U), p(a), p(b), p(q), p(sh), C.c_int64(m), C.byref(Cfg(**cfg)), p(lw))
        return lw

    def lse(lw):
        mx = lw.max()
        return mx + np.log(np.mean(np.exp(lw - mx)))

    base = run(FP64)
    r = O.qmc_pnorm(F.U, F.a, F.b, m=m, n_reps=1, shifts=P.shifts[:, :1], phi_form="tail", bs_i=64, bs_i2=32)
    print(f"factor model d={d}, {m} points of one shift; all-FP64 emulation vs oracle: "
          f"|d log P| = {abs(lse(base) - r.log_value):.1e}")
    print("| stage in FP32 (the rest FP64) | |d log P| | per-sample mean | per-sample std |")
    print("|---|---|---|---|")

    def one(**kw):
        c = dict(FP64)
        c.update(kw)
        return c

    rows = [("points (32-bit fixed point)", one(pt32=1)),
            ("panel contraction operands (split-FP16, 22 bits)", one(pbits=22)),
            ("panel contraction operands (three FP16 parts, 33 bits)", one(pbits=33)),
            ("panel contraction accumulation (FP32)", one(pacc=1)),
            ("in-panel update (FP32 matrix cores)", one(inpanel=1)),
            ("in-block dot (FP32)", one(inblock=1)),
            ("z = (b - s) inv (FP32)", one(z=1)),
            ("z exact, rounded to FP32 once", one(z=2)),
            ("tail mass (FP32 fit)", one(mass=1, state=1)),
            ("quantile (FP32 fit)", one(quant=1, state=1)),
            ("weight product (FP32)", one(weight=1, state=1)),
            ("all (the device's mixed walk)", dict(DEVICE))]
    dev = None
    for name, cfg in rows:
        lw = run(cfg)
        if cfg == DEVICE:
            dev = lw
        dl = lw - base
        print(f"| {name} | {abs(lse(lw) - lse(base)):.1e} | {dl.mean():+.1e} | {dl.std():.1e} |")
    o = np.argsort(-base)
    w = np.exp(base[o] - base.max())
    cum = np.cumsum(w) / w.sum()
    print(f"\nlog weights over the samples: std {base.std():.1f}; the heaviest sample holds "
          f"{w[0] / w.sum():.6f} of the sum, the heaviest {int(np.searchsorted(cum, 1 - 1e-8)) + 1} all but 1e-8")
    od = np.argsort(-dev)
    for k in (1, 4, 16):
        lw = dev.copy()
        lw[od[:k]] = base[od[:k]]
        print(f"device walk with its {k} heaviest samples in FP64: |d log P| = {abs(lse(lw) - lse(base)):.1e}")


if __name__ == "__main__":
    main()
