// Lower-tail normal quantile in FP32 for a probability mass given as
// p = t * 2^-k (t > 0 a normal FP32 number, k >= 0 an integer scale), the
// form in which the mixed-precision integration carries masses that fall
// below the FP32 range.  Returns Phi^{-1}(p) for p <= 1/2.
//
// Method: Wichura's algorithm AS 241 (Applied Statistics 37, 1988), the
// PPND16 rational approximations rounded to FP32:
//   |p - 1/2| <= 0.425:  q (A(r) / B(r)),  r = 0.180625 - q^2, q = p - 1/2;
//   otherwise:           r = sqrt(-log p), C(r - 1.6) / D(r - 1.6) for r <= 5,
//                        E(r - 5) / F(r - 5) beyond.
// -log p is formed as -log t + k log 2 in FP32, so masses far below the FP32
// range need no FP64 arithmetic.  The result is accurate to a few FP32 ulps
// (the rationals themselves are good to about 1e-16).
//
// The same source compiles for the host (plain C++) so that the CPU test
// suite can check it against a high-precision quantile.
#pragma once

#include <math.h>

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
// 2^-1300, beyond the FP64 range too); the caller handles infinite z.
//
// Method: erfc(z) = t exp(-z^2 + P(t)), t = 2 / (2 + z) (the form of
// Numerical Recipes' erfcc, Press et al. 1992, sec. 6.2), with P a degree-11
// polynomial in u = 2t - 1 = (2 - z) / (2 + z) fitted by least squares on 600
// Chebyshev nodes of t in [0, 1] to 50-digit values of
// log(erfc(z) / t) + z^2 (fit error 1.1e-8 over all z >= 0; the script that
// made it is fit_tail_mass.py beside this file).  u is formed from z directly
// (a refined quotient), not from a rounded t, since P is steep in u.  z^2 is
// split exactly as p + e (e = fma(z, z, -p)) and the exponential reduced
// Cody-Waite style: n = rint((P - p) / ln2),
// r = ((-p + n ln2_hi) + P - e) + n ln2_lo (the first sum is exact), so that
// erfc(z)/2 = 2^n exp(r) / (2 + z).  Relative error about 2.5e-7 with exact
// reciprocal and exp2, 5e-7 with the device's 1-ulp v_rcp_f32 and v_exp_f32.
DBG_HD inline float tail_mass_f32(float z, int& k)
{
    z = fminf(z, 30.0f);
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

DBG_HD inline float tail_quantile_f32(float t, int k)
{
    const float p = ldexpf(t, -k);
    const float q = p - 0.5f;
    if (fabsf(q) <= 0.425f) {
        const float r = fmaf(-q, q, 0.180625f);
        const float A[8] = {2.5090809287301226727e+3f, 3.3430575583588128105e+4f, 6.7265770927008700853e+4f,
                            4.5921953931549871457e+4f, 1.3731693765509461125e+4f, 1.9715909503065514427e+3f,
                            1.3314166789178437745e+2f, 3.3871328727963666080e+0f};
        const float B[8] = {5.2264952788528545610e+3f, 2.8729085735721942674e+4f, 3.9307895800092710610e+4f,
                            2.1213794301586595867e+4f, 5.3941960214247511077e+3f, 6.8718700749205790830e+2f,
                            4.2313330701600911252e+1f, 1.0f};
        return q * div_pos(horner_f32(r, A), horner_f32(r, B));
    }
    // -log p = -log t + k log 2
    const float s = fmaf((float)k, 0.693147180559945309f, -logf(t));
    float r = sqrtf(s);
    float val;
    if (r <= 5.0f) {
        const float C[8] = {7.74545014278341407640e-4f, 2.27238449892691845833e-2f, 2.41780725177450611770e-1f,
                            1.27045825245236838258e+0f, 3.64784832476320460504e+0f, 5.76949722146069140550e+0f,
                            4.63033784615654529590e+0f, 1.42343711074968357734e+0f};
        const float D[8] = {1.05075007164441684324e-9f, 5.47593808499534494600e-4f, 1.51986665636164571966e-2f,
                            1.48103976427480074590e-1f, 6.89767334985100004550e-1f, 1.67638483018380384940e+0f,
                            2.05319162663775882187e+0f, 1.0f};
        r -= 1.6f;
        val = div_pos(horner_f32(r, C), horner_f32(r, D));
    } else {
        const float E[8] = {2.01033439929228813265e-7f, 2.71155556874348757815e-5f, 1.24266094738807843860e-3f,
                            2.65321895265761230930e-2f, 2.96560571828504891230e-1f, 1.78482653991729133580e+0f,
                            5.46378491116411436990e+0f, 6.65790464350110377720e+0f};
        const float F[8] = {2.04426310338993978564e-15f, 1.42151175831644588870e-7f, 1.84631831751005468180e-5f,
                            7.86869131145613259100e-4f, 1.48753612908506148525e-2f, 1.36929880922735805310e-1f,
                            5.99832206555887937690e-1f, 1.0f};
        r -= 5.0f;
        val = div_pos(horner_f32(r, E), horner_f32(r, F));
    }
    return -val;
}

}  // namespace dbg
