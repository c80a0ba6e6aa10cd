"""Host-side checks that need no GPU: the C ABI library loads and exports
every function include/dbgenz.h declares; option defaults, constructor
clamps, shift drawing and the synthetic workloads."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    src = open(os.path.join(ROOT, "include", "dbgenz.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dbg_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(dbg):
    from dbgenzpaper_amd import _lib

    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared_functions()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.EXPORTS) == set(names)


def test_abi_version_and_defaults(dbg):
    from dbgenzpaper_amd import _lib

    L = _lib.lib()
    assert L.dbg_abi_version() == 3
    o = _lib.Opts()
    # the ctypes mirror has the size of the C struct (int32/int64 fields, 8-byte aligned)
    assert ctypes.sizeof(o) == 96
    L.dbg_opts_default(ctypes.byref(o), _lib.PREC_FP64)
    assert o.contraction == _lib.CONTRACT_AUTO and o.refine_tiles == -1
    # QMC_opts defaults, src/qmc.jl:111-125
    assert (o.chol_block_size, o.chol_block_size2, o.m, o.block_size_i, o.block_size_j,
            o.block_size_i2, o.n_reps, o.n_bits, o.max_pts) == (64, 512, 10000, 64, 64, 32, 12, 32, 10000)
    assert o.max_abs_err == 1e-6 and o.sobol_skip == -1


def test_qmc_opts_mirror(dbg):
    o = dbg.QMC_opts()
    assert (o.chol_block_size, o.chol_block_size2, o.m, o.block_size_i, o.block_size_i2,
            o.block_size_j, o.n_reps, o.n_bits, o.max_pts, o.max_abs_err) == \
           (2 ** 6, 2 ** 9, 10 ** 4, 2 ** 6, 2 ** 5, 2 ** 6, 12, 32, 10 ** 4, 1e-6)
    assert dbg.QMC_opts(np.float32).T == np.float32
    with pytest.raises(TypeError):
        dbg.QMC_opts(np.int32)


def test_constructor_clamps():
    from dbgenzpaper_amd import _lib as L
    from dbgenzpaper_amd.qmc import QMC_opts, _clamped_opts

    o = QMC_opts(block_size_i=64, block_size_i2=32)
    c = _clamped_opts(o, 10, L.KIND_DENSE)          # src/qmc.jl:166-174
    assert (c.block_size_i, c.block_size_i2) == (9, 9)
    c = _clamped_opts(o, 40, L.KIND_DENSE)
    assert (c.block_size_i, c.block_size_i2) == (39, 32)
    c = _clamped_opts(o, 10, L.KIND_SPARSE)         # src/qmc.jl:239-242
    assert c.block_size_i == 10


def test_draw_shifts_shapes(dbg):
    s = dbg.draw_shifts(np.random.default_rng(0), 9, 5, "Richtmyer")
    assert s.shape == (8, 5) and s.dtype == np.float64 and np.all((s >= 0) & (s < 1))
    u = dbg.draw_shifts(np.random.default_rng(0), 9, 5, "Sobol")
    assert u.shape == (8, 5) and u.dtype == np.uint64
    f = dbg.draw_shifts(np.random.default_rng(0), 9, 5, "Richtmyer", np.float32)
    assert np.all(f == f.astype(np.float32))


def test_workloads():
    from scipy.special import ndtr

    from dbgenzpaper_amd import workloads as W

    P = W.factor_model(50, 8, m=16, n_reps=3)
    assert np.allclose(np.diag(P.Sigma), 1) and np.all(np.linalg.eigvalsh(P.Sigma) > 0)
    u = ndtr(P.b)
    assert np.all((u > 0.0099) & (u < 0.1001)) and np.all(np.isneginf(P.a))
    B = W.banded(64, 4, m=16, n_reps=2)
    U = np.linalg.cholesky(B.Sigma).T
    assert np.allclose(np.diag(B.Sigma), 1) and np.abs(np.triu(U, 5)).max() < 1e-12
    batch = W.batch_uneven(12, m=16)
    assert len(batch) == 12 and all(2 <= p.d <= 256 for p in batch)


def test_library_missing_fails_loudly(tmp_path, monkeypatch):
    from dbgenzpaper_amd import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libdbgenz.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.lib()


def _host_build(tmp_path, src, name):
    """Compile a small host program against the device headers (plain C++)."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    csrc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "dbgenzpaper_amd", "csrc")
    cpp = tmp_path / f"{name}.cpp"
    cpp.write_text(src)
    so = tmp_path / f"lib{name}.so"
    subprocess.run([gxx, "-O2", "-ffp-contract=off", "-shared", "-fPIC", "-I", csrc, str(cpp), "-o", str(so)],
                   check=True)
    return ctypes.CDLL(str(so))


@pytest.mark.parametrize("rcp_ulp", [0, 1, -1])
def test_fp32_tail_quantile_against_mpmath(tmp_path, rcp_ulp):
    """The mixed path's FP32 inverse CDF (two fitted polynomials in
    w = -log(4p(1-p)), w formed from t and k) against a 40-digit quantile,
    from p = 0.5 down to p = 2^-3000 given as scaled masses t 2^-k.  Bound:
    4 FP32 ulps of max(|y|, 1).  The tail region takes 1/sqrt(w) from a
    reciprocal square root; rcp_ulp = +-1 builds it one ulp off, as the
    device's approximate v_rsq_f32 may be."""
    mpmath = pytest.importorskip("mpmath")
    pre = f"#define DBG_RCP_ULP_TEST {rcp_ulp}\n" if rcp_ulp else ""
    lib = _host_build(tmp_path, pre + '#include "tailq.h"\n'
                      'extern "C" float tq(float t, int k) { return dbg::tail_quantile_f32(t, k); }\n', "tq")
    lib.tq.restype = ctypes.c_float
    lib.tq.argtypes = [ctypes.c_float, ctypes.c_int]
    mpmath.mp.dps = 40

    def truth(p):
        y = mpmath.mpf(-0.5) if p > 0.1 else -mpmath.sqrt(-2 * mpmath.log(p))
        for _ in range(200):
            f = mpmath.log(mpmath.ncdf(y)) - mpmath.log(p)
            yn = y - f * mpmath.ncdf(y) / mpmath.npdf(y)
            if abs(yn - y) < mpmath.mpf(10) ** -30:
                return yn
            y = yn
        return y

    worst = 0.0
    # plus both sides of the region switch, w = 6.25 (p = 4.83e-4), and p near 5.80e-17 (sqrt(w) = 6)
    edges = np.log2(np.concatenate([4.8303e-4 * (1 + np.linspace(-1e-3, 1e-3, 21)),
                                    5.7988e-17 * (1 + np.linspace(-1e-3, 1e-3, 21))]))
    log2p = np.concatenate([np.linspace(-1.0, -40.0, 157), np.linspace(-40.0, -3000.0, 97), edges])
    for lp in log2p:
        k = max(0, int(-lp) - 100) if lp < -110 else 0
        t = float(np.float32(2.0 ** (lp + k)))
        p = mpmath.mpf(t) * mpmath.mpf(2) ** (-k)
        tr = truth(p)
        y = lib.tq(t, k)
        worst = max(worst, float(abs(y - tr) / max(abs(tr), 1)))
    # p = 1/2 gives 0 exactly
    assert lib.tq(0.5, 0) == 0.0
    assert worst < 4 * np.finfo(np.float32).eps, worst


@pytest.mark.parametrize("rcp_ulp", [0, 1, -1])
def test_fp32_tail_mass_against_mpmath(tmp_path, rcp_ulp):
    """The mixed path's FP32 tail mass erfc(z)/2 = m 2^-k against 40-digit
    erfc over z in [0, 30] (masses down to 2^-1300), with the reciprocal exact
    and one ulp off either way (the device's v_rcp_f32).  Bound 5e-7 relative
    (about 8 FP32 ulps); m stays in [0.02, 0.71] and z is capped at 30."""
    mpmath = pytest.importorskip("mpmath")
    pre = f"#define DBG_RCP_ULP_TEST {rcp_ulp}\n" if rcp_ulp else ""
    lib = _host_build(tmp_path, pre + '#include "tailq.h"\n'
                      'extern "C" float tm(float z, int* k) { return dbg::tail_mass_f32(z, *k); }\n', "tm")
    lib.tm.restype = ctypes.c_float
    lib.tm.argtypes = [ctypes.c_float, ctypes.POINTER(ctypes.c_int)]
    mpmath.mp.dps = 40
    k = ctypes.c_int()
    worst, lo, hi = 0.0, 1.0, 0.0
    zs = np.concatenate([np.linspace(0.0, 6.0, 1201), np.linspace(6.0, 30.0, 481), [1e-30, 2.0, 9.0, 26.5]])
    for z in zs.astype(np.float32):
        m = lib.tm(float(z), ctypes.byref(k))
        assert k.value >= 0
        lo, hi = min(lo, m), max(hi, m)
        tr = mpmath.erfc(mpmath.mpf(float(z))) / 2
        worst = max(worst, float(abs(mpmath.mpf(m) * mpmath.mpf(2) ** (-k.value) / tr - 1)))
    assert worst < 5e-7, worst
    assert 0.02 <= lo and hi <= 0.71, (lo, hi)
    kc = ctypes.c_int()
    assert lib.tm(45.0, ctypes.byref(k)) == lib.tm(30.0, ctypes.byref(kc)) and k.value == kc.value


def test_tent_and_plain_maps_exact(tmp_path):
    """The walk's map from a point's fixed-point coordinate u (32 fraction
    bits) to FP32 w and 1 - w: the tent map w = |2u - 1|
    (src/qmc_generators.jl:160) and, for the SparseQMC module, w = u, each of
    w and 1 - w the correctly rounded FP32 value of the exact fraction.  Exact
    values from integer arithmetic (|2u - 2^32| and its complement are
    integers below 2^33, exact in FP64 and scaled by 2^-32 exactly, then
    rounded once to FP32), on every u with 18 leading bits fixed at random
    plus the edges 0, 1/2 and the values next to them."""
    lib = _host_build(tmp_path, '#include "tailq.h"\n'
                      'extern "C" void maps(const unsigned* u, long n, float* tw, float* tc, float* pw, float* pc) {\n'
                      '  for (long i = 0; i < n; ++i) { dbg::tent_w(u[i], tw[i], tc[i]); dbg::plain_w(u[i], pw[i], pc[i]); }\n'
                      '}\n', "maps")
    rng = np.random.default_rng(11)
    lo = np.arange(1 << 14, dtype=np.uint64)
    hi = rng.integers(0, 1 << 18, 64, dtype=np.uint64) << np.uint64(14)
    edges = np.array([0, 1, 2, 3, 2**31 - 2, 2**31 - 1, 2**31, 2**31 + 1, 2**31 + 2, 2**32 - 2, 2**32 - 1],
                     dtype=np.uint64)
    u = np.concatenate([(hi[:, None] + lo[None, :]).ravel(), edges]).astype(np.uint32)
    n = u.size
    tw, tc, pw, pc = (np.empty(n, np.float32) for _ in range(4))
    f = ctypes.POINTER(ctypes.c_float)
    lib.maps(u.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), ctypes.c_long(n), tw.ctypes.data_as(f),
             tc.ctypes.data_as(f), pw.ctypes.data_as(f), pc.ctypes.data_as(f))
    ui = u.astype(np.int64)
    a = np.abs(2 * ui - 2**32)  # 2^32 |2u - 1|
    s = 2.0 ** -32
    np.testing.assert_array_equal(tw, (a * s).astype(np.float32))
    np.testing.assert_array_equal(tc, ((2**32 - a) * s).astype(np.float32))
    np.testing.assert_array_equal(pw, (ui * s).astype(np.float32))
    np.testing.assert_array_equal(pc, ((2**32 - ui) * s).astype(np.float32))
    # u = 0 and u = 1/2: w = 1 and 0
    assert (tw[-11], tc[-11], tw[-5], tc[-5]) == (1.0, 0.0, 0.0, 1.0)


def test_plan_option_validation(dbg):
    """dbg_plan_create rejects invalid options before touching the device:
    the contraction option (DBG_CONTRACT_*) is checked against the path."""
    from dbgenzpaper_amd import _lib

    L = _lib.lib()
    n = 8
    S = np.asfortranarray(np.eye(n))
    a, b = -np.ones(n), np.ones(n)
    h = ctypes.c_void_p()

    def create(prec, **fields):
        o = _lib.Opts()
        L.dbg_opts_default(ctypes.byref(o), prec)
        for k, v in fields.items():
            setattr(o, k, v)
        return L.dbg_plan_create(ctypes.byref(h), ctypes.byref(o), n, _lib.ptr(S, ctypes.c_double), n,
                                 _lib.ptr(a, ctypes.c_double), _lib.ptr(b, ctypes.c_double))

    assert create(_lib.PREC_MIXED, contraction=7) == _lib.DBG_ERR_ARG
    assert b"unknown contraction" in L.dbg_last_error()
    assert create(_lib.PREC_FP64, contraction=_lib.CONTRACT_SPLIT16) == _lib.DBG_ERR_ARG
    assert b"split-FP16" in L.dbg_last_error()
    assert create(_lib.PREC_MIXED, contraction=_lib.CONTRACT_SPLIT16, kind=_lib.KIND_SPARSE) == _lib.DBG_ERR_ARG
    assert create(_lib.PREC_MIXED, n_bits=0) == _lib.DBG_ERR_ARG
    b[3] = np.nan
    assert create(_lib.PREC_MIXED) == _lib.DBG_ERR_ARG and b"NaN" in L.dbg_last_error()
    b[3] = 1.0
    assert create(_lib.PREC_MIXED, m=0) == _lib.DBG_ERR_ARG
    assert not h.value


def test_cov_build_validation(dbg):
    """dbg_cov_build rejects invalid generators before touching the device."""
    from dbgenzpaper_amd import _lib

    L = _lib.lib()
    n = 8
    S = np.zeros((n, n), order="F")

    def build(kind, params, k=0, B=None, psi=None, lds=n):
        p = np.array(params, np.float64)
        return L.dbg_cov_build(kind, n, _lib.ptr(p, ctypes.c_double), k, B, psi, 0, S.ctypes.data, lds, 0)

    assert build(7, [1, 0, 0]) == _lib.DBG_ERR_ARG
    assert build(_lib.COV_MATERN, [1.0, 1.0, 1.0]) == _lib.DBG_ERR_ARG      # nu not half-integer
    assert b"nu must be" in L.dbg_last_error()
    assert build(_lib.COV_MATERN, [1.5, 0.0, 1.0]) == _lib.DBG_ERR_ARG      # dist = 0
    assert build(_lib.COV_FIXED_SPARSE, [2.5, 0, 0]) == _lib.DBG_ERR_ARG    # fractional bandwidth
    assert build(_lib.COV_LOWRANK, [0, 0, 0]) == _lib.DBG_ERR_ARG           # B and psi missing
    assert build(_lib.COV_FIXED_DENSE, [1, 0, 0], lds=n - 1) == _lib.DBG_ERR_ARG


@pytest.mark.parametrize("nu", [0.5, 1.5, 2.5])
def test_matern_closed_forms(nu):
    """The half-integer closed forms the device generator evaluates equal the
    reference's general formula sigma2 2^(1-nu)/Gamma(nu) z^nu K_nu(z)
    (scripts/simulations/simulations_functions.jl:56-79) through SciPy's kv."""
    from scipy.special import gamma, kv

    z = np.concatenate([np.linspace(1e-3, 5, 400), np.linspace(5, 40, 200)])
    general = 2 ** (1 - nu) / gamma(nu) * z ** nu * kv(nu, z)
    closed = {0.5: np.exp(-z), 1.5: (1 + z) * np.exp(-z), 2.5: (1 + z + z * z / 3) * np.exp(-z)}[nu]
    assert np.allclose(closed, general, rtol=1e-13, atol=0)


_CTYPE = {"int32_t": "c_int", "int64_t": "c_long", "double": "c_double"}
_JULIA = {"Int32": "c_int", "Int64": "c_long", "Float64": "c_double"}


def _c_struct_fields(src, name):
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), src, flags=re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    return [(f, _CTYPE[t]) for t, f in re.findall(r"(int32_t|int64_t|double)\s+(\w+);", body)]


def _julia_struct_fields(md, name):
    body = re.search(r"struct %s\b[^\n]*\n(.*?)\nend" % name, md, flags=re.S).group(1)
    body = re.sub(r"#[^\n]*", "", body)
    return [(f, _JULIA[t]) for f, t in re.findall(r"(\w+)::(Int32|Int64|Float64)", body)]


def test_julia_binding_matches_the_header():
    """The Julia structs of the binding guide at the repository root have the
    C structs' fields in order and type, the ctypes mirror too, and the
    guide's routing function builds DbgOpts with one argument per field."""
    from dbgenzpaper_amd import _lib

    hdr = open(os.path.join(ROOT, "include", "dbgenz.h")).read()
    md = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    for cname, jname, mirror in (("dbg_opts", "DbgOpts", _lib.Opts), ("dbg_result", "DbgResult", _lib.Result)):
        c = _c_struct_fields(hdr, cname)
        j = _julia_struct_fields(md, jname)
        py = [(f, t.__name__) for f, t in mirror._fields_]
        py = [(f, {"c_int32": "c_int", "c_int64": "c_long"}.get(t, t)) for f, t in py]
        assert c == j, (cname, c, j)
        assert c == py, (cname, c, py)
    # DbgOpts(...) in the routing function: one argument per field
    call = md[md.index("o = DbgOpts(") + len("o = DbgOpts("):]
    depth, args, cur = 1, [], ""
    for ch in call:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
            if depth == 0:
                args.append(cur)
                break
        if ch == "," and depth == 1:
            args.append(cur)
            cur = ""
        else:
            cur += ch
    assert len(args) == len(_c_struct_fields(hdr, "dbg_opts")), args
