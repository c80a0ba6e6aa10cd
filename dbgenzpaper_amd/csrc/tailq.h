// FP32 normal tail functions of the mixed-precision integration, which
// carries probability masses as p = t 2^-k (t a normal FP32 number, k >= 0 an
// integer scale) so that masses far below the FP32 range need no FP64:
//   tail_mass_f32      erfc(z)/2 for z >= 0, as m 2^-k;
//   tail_quantile_f32  Phi^{-1}(p) for p = t 2^-k <= 1/2;
// and the map from a point's fixed-point coordinate u (32 fraction bits) to
// the FP32 pair (w, 1 - w) that the walk samples with:
//   tent_w             w = |2u - 1| (the tent map);
//   plain_w            w = u (no tent map: the SparseQMC module).
// The two tail functions are polynomials fitted for this code (fit_tailq.py
// beside this file reproduces every coefficient) and are branch-free or nearly
// so, since the integration evaluates them once per sample and dimension in
// SIMD.
//
// The same source compiles for the host (plain C++) so that the CPU test
// suite can check these functions against exact or high-precision values.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define DBG_HD __host__ __device__
#else
#define DBG_HD
#endif

namespace dbg {

// 1/d: the device's approximate v_rcp_f32 (within 1 ulp), the host's division
DBG_HD inline float rcp_f32(float d)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(d);
#elif defined(DBG_RCP_ULP_TEST)
    // host check of the refinements: the reciprocal off by one ulp
    return nextafterf(1.0f / d, DBG_RCP_ULP_TEST > 0 ? INFINITY : 0.0f);
#else
    return 1.0f / d;
#endif
}

// n / d for the positive, well-scaled denominators of the rational
// approximations below: a reciprocal (the device's approximate v_rcp_f32,
// within 1 ulp; the host's division) refined by one Newton step on the
// quotient, q = q0 + (n - d q0) / d, which leaves the result within about one
// ulp of the correctly rounded quotient at a fraction of the cost of a full
// IEEE division (no special cases: d is never 0, inf or subnormal here).
DBG_HD inline float div_pos(float n, float d)
{
    const float r = rcp_f32(d);
    const float q0 = n * r;
    return fmaf(fmaf(-d, q0, n), r, q0);
}

// Horner evaluation with fused multiply-adds, highest coefficient first
template <int N>
DBG_HD inline float horner_f32(float x, const float (&c)[N])
{
    float v = c[0];
#pragma unroll
    for (int i = 1; i < N; ++i) v = fmaf(v, x, c[i]);
    return v;
}

// Upper normal tail mass erfc(z)/2 (= 1 - Phi(z sqrt2)) for z >= 0 in FP32,
// returned as m 2^-k with m in [0.02, 0.71] and k >= 0, so that it stays
// representable far past the FP32 range.  z is capped at 30 (a mass near
// 2^-1300, beyond the FP64 range too); the caller handles infinite z, and z is
// never NaN (finite factor, bounds checked for NaN when the plan is made).
//
// Method: erfc(z) = t exp(-z^2 + P(t)), t = 2 / (2 + z) (the form of
// Numerical Recipes' erfcc, Press et al. 1992, sec. 6.2), with P a degree-11
// polynomial in u = 2t - 1 = (2 - z) / (2 + z) fitted by least squares on 600
// Chebyshev nodes of t in [0, 1] to 50-digit values of
// log(erfc(z) / t) + z^2 (fit error 1.1e-8 over all z >= 0; the script that
// made it is fit_tailq.py beside this file).  u is formed from z directly
// (a refined quotient), not from a rounded t, since P is steep in u.  z^2 is
// split exactly as p + e (e = fma(z, z, -p)) and the exponential reduced
// Cody-Waite style: n = rint((P - p) / ln2),
// r = ((-p + n ln2_hi) + P - e) + n ln2_lo (the first sum is exact), so that
// erfc(z)/2 = 2^n exp(r) / (2 + z).  Relative error about 2.5e-7 with exact
// reciprocal and exp2, 5e-7 with the device's 1-ulp v_rcp_f32 and v_exp_f32.
DBG_HD inline float tail_mass_f32(float z, int& k)
{
    z = fminf(z, 30.0f);  // one v_min_f32; NaN bounds are rejected where a plan is created
    const float s = 2.0f + z;
    const float h = rcp_f32(s);
    const float w = 2.0f - z;
    const float q0 = w * h;
    const float u = fmaf(fmaf(-s, q0, w), h, q0);
    const float C[12] = {-8.728381363e-05f, 8.000738489e-06f, 5.737663596e-04f, -2.279109613e-04f,
                         -2.288765274e-03f, 1.794066280e-03f, 8.808825165e-03f, -9.880019352e-03f,
                         -4.689364135e-02f, 4.734388739e-02f, 6.726431251e-01f, -6.717941165e-01f};
    const float P = horner_f32(u, C);
    const float p = z * z;
    const float e = fmaf(z, z, -p);
    const float n = rintf((P - p) * 1.44269504088896341f);
    float r = fmaf(-n, 0.693145751953125f, -p);  // ln2_hi: 16 bits, the sum is exact
    r = r + P;
    r = fmaf(-n, 1.428606765330187e-06f, r - e);  // ln2_lo
#if defined(__HIP_DEVICE_COMPILE__)
    const float ex = __builtin_amdgcn_exp2f(r * 1.44269504088896341f);
#else
    const float ex = exp2f(r * 1.44269504088896341f);
#endif
    k = -(int)n;
    return h * ex;
}

DBG_HD inline float log2_f32(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_logf(x);
#else
    return log2f(x);
#endif
}

// 1/sqrt(x) for the positive, normal arguments of tail_quantile_f32: the
// device's approximate v_rsq_f32 (within 1 ulp), the host's correctly rounded
// quotient (or one ulp off it, for the host check of that tolerance)
DBG_HD inline float rsqrt_f32(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#elif defined(DBG_RCP_ULP_TEST)
    return nextafterf(1.0f / sqrtf(x), DBG_RCP_ULP_TEST > 0 ? INFINITY : 0.0f);
#else
    return 1.0f / sqrtf(x);
#endif
}

// Phi^{-1}(p), p = t 2^-k <= 1/2, in terms of w = -log(4p(1 - p)) (the
// variable of Giles, "Approximating the erfinv function", GPU Computing Gems
// 2010; coefficients fitted here) in two regions: y / (2p - 1) is a
// polynomial in w for w < 6.25 (p > 4.8e-4, the common case at moderate
// dimension), and beyond it -y / s is a polynomial in x = 1/s, s = sqrt(w),
// fitted for s in [2.5, 50] (p from 4.8e-4 down to about 2^-3600), with x from
// one reciprocal square root and s = w x refined by one Newton step (the
// factor s carries its rounding into y; x enters only through the flat
// polynomial).  w is formed as
// (k - log2(4t(1 - p))) log 2, so no step leaves the FP32 range.  Within 1.7
// FP32 eps of max(|y|, 1) on the host, with the reciprocal square root exact
// or one ulp off.
DBG_HD inline float tail_quantile_f32(float t, int k)
{
    const float p = ldexpf(t, -k);  // 0 when far below the FP32 range
    // w = -log(4 p (1 - p)) = (k - log2(4 t (1 - p))) log 2
    const float w = (float(k) - log2_f32(4.0f * t * (1.0f - p))) * 0.693147180559945309f;
    if (w < 6.25f) {  // p > 4.8e-4
        const float C1[11] = {9.308407933e-10f, -6.508231731e-09f, -3.488968403e-08f, 6.018015029e-07f,
                              -1.963471277e-06f, -1.963790237e-05f, 2.641592582e-04f, -1.047509722e-03f,
                              -8.532967418e-03f, 3.396349549e-01f, 2.338620663e+00f};
        return fmaf(2.0f, p, -1.0f) * horner_f32(w - 3.125f, C1);
    }
    const float x = rsqrt_f32(w);
    const float s0 = w * x;
    const float s = fmaf(0.5f * x, fmaf(-s0, s0, w), s0);  // one Newton step: s to about half an ulp
    const float CB[13] = {-1.301323203e+05f, 1.896112891e+04f, 1.034813477e+04f, -9.892526855e+02f,
                          -4.395959778e+02f, 4.312902832e+01f, 4.628243446e-01f, 2.000564814e+00f,
                          -1.639978766e+00f, 1.008415103e+00f, 6.785751786e-03f, -2.852775156e-01f,
                          1.368908286e+00f};
    return -s * horner_f32(x - 0.21f, CB);
}

// w = |2u - 1| (the tent map of src/qmc_generators.jl:160) and 1 - w in FP32
// from the fixed-point u; u = 0 gives w = 1.
DBG_HD inline void tent_w(uint32_t u, float& w, float& wc)
{
    // with u2 = 2u mod 2^32: for u >= 1/2, w = u2 2^-32 and 1 - w = (2^32 - u2)
    // 2^-32; for u < 1/2 the two swap.  2^32 - u2 is -u2 mod 2^32 except at
    // u2 = 0 (u = 0 or 1/2), where it is 2^32: the value 1.
    const uint32_t u2 = u << 1;
    constexpr float kS = 2.3283064365386963e-10f;  // 2^-32
    const float a = (float)u2 * kS;
    const float b = (u2 == 0u) ? 1.0f : (float)(0u - u2) * kS;
    const bool hi = (int32_t)u < 0;
    w = hi ? a : b;
    wc = hi ? b : a;
}

// w = u and 1 - u without the tent map (the SparseQMC module)
DBG_HD inline void plain_w(uint32_t u, float& w, float& wc)
{
    constexpr float kS = 2.3283064365386963e-10f;
    w = (float)u * kS;
    wc = (u == 0u) ? 1.0f : (float)(0u - u) * kS;
}

}  // namespace dbg
