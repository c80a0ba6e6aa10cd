// Device-side building blocks shared by the integration kernels: the fused
// Richtmyer / Sobol point map, the two per-dimension precisions (reference
// FP64 arithmetic and the FP32 mixed form), the scaled weight product and
// the block reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dbg_internal.h"

namespace dbg {

constexpr double kSqrt2 = 1.4142135623730951;       // Float64(sqrt(2.0))
constexpr double kEps64 = 2.220446049250313e-16;    // eps(Float64)

// ---------------------------------------------------------------------------
// Point generators (src/qmc_generators.jl)
// ---------------------------------------------------------------------------

// frac(Float64(n0 + j) * q)   richtmyer_mat!  src/qmc_generators.jl:314-325
__device__ __forceinline__ double richtmyer_x(int64_t n0_plus_j, double q)
{
    const double x = __dmul_rn((double)n0_plus_j, q);
    return __dsub_rn(x, floor(x));
}

// Sobol integer of sequence index idx: XOR of the direction integers
// selected by the Gray code of idx (equals floor(x * 2^32) of
// copyt_as_UInt64!, src/qmc_generators.jl:340-352).  V points at the 32
// direction integers of one dimension (uniform across the wavefront).
__device__ __forceinline__ uint32_t sobol_x(const uint32_t* __restrict__ V, uint64_t idx)
{
    const uint64_t g = idx ^ (idx >> 1);
    const uint32_t glo = (uint32_t)g;
    uint32_t x = 0;
#pragma unroll
    for (int k = 0; k < 32; ++k) x ^= ((glo >> k) & 1u) ? V[k] : 0u;
    return x;
}

// Uniform u in [0, 1] of one (point, dimension, shift) before the tent map:
// Richtmyer shift mod 1 (src/qmc_generators.jl:158-159) or Sobol digital shift
// (:197-198).  Computed in FP64 for both precisions.
template <int GEN>
__device__ __forceinline__ double point_u(const PointArgs& P, int g, int k, int64_t j)
{
    if constexpr (GEN == kGenRichtmyer) {
        const double shift = P.shifts_f[(size_t)g + (size_t)k * P.nd];
        double u = __dadd_rn(shift, richtmyer_x(P.n_base + j, P.q[g]));
        return u > 1.0 ? __dsub_rn(u, 1.0) : u;
    } else {
        const uint64_t mask = P.mask;
        const uint64_t shift = P.shifts_u[(size_t)g + (size_t)k * P.nd] & mask;
        const uint64_t x = (uint64_t)sobol_x(P.V + (size_t)g * 32, P.sobol_first + (uint64_t)j) & mask;
        return __dmul_rn((double)((x ^ shift) & mask), P.invscale);
    }
}

// ---------------------------------------------------------------------------
// Scaled weight product: p = pm * 2^pe with pm in [0.5, 1) (or 0).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double frexp_(double x, int* e) { return frexp(x, e); }
__device__ __forceinline__ float frexp_(float x, int* e) { return frexpf(x, e); }

template <typename T>
struct Weight {
    T pm;
    int pe;
    __device__ __forceinline__ void init(T dc)
    {
        int e = 0;
        pm = frexp_(dc, &e);
        pe = (pm == T(0)) ? 0 : e;
    }
    __device__ __forceinline__ void mul(T dc)
    {
        int e = 0;
        pm = frexp_(pm * dc, &e);
        pe += (pm == T(0)) ? 0 : e;
    }
};

// ---------------------------------------------------------------------------
// FP64 per-dimension state: reference arithmetic of qmc_loop!.
// ---------------------------------------------------------------------------
struct StateF64 {
    double c, dc;
};

// rand_points! + Y = sqrt2 * erfinv(Y)   (src/qmc_generators.jl:156-164,
// src/qmc.jl:385-387).  u is the shifted uniform of point_u.
// sample_f64_w takes the already tent-mapped point w (the SparseQMC module
// has no tent map: src/qmc_sparse.jl:202-207).
__device__ __forceinline__ double sample_f64_w(const StateF64& st, double w)
{
    double u = fma(w, st.dc, st.c);
    u = fmin(fmax(u, kEps64), 1.0 - kEps64);
    const double Yv = fma(2.0, u, -1.0);
    return __dmul_rn(kSqrt2, erfinv(Yv));
}

__device__ __forceinline__ double sample_f64(const StateF64& st, double u)
{
    return sample_f64_w(st, fabs(fma(u, 2.0, -1.0)));
}

// c, dc of the next dimension from s  (src/qmc.jl:405-429 erf form,
// :579-593 erfc form).
template <int FORM>
__device__ __forceinline__ void phi_f64(StateF64& st, double s, double a, double b, double inv)
{
    double x = __dmul_rn(__dsub_rn(a, s), inv);
    double xb = __dmul_rn(__dsub_rn(b, s), inv);
    x = fmax(fmin(x, 9.0), -9.0);
    xb = fmax(fmin(xb, 9.0), -9.0);
    if constexpr (FORM == kPhiErf) {
        const double e1 = erf(x), e2 = erf(xb);
        st.c = __dmul_rn(0.5, __dadd_rn(1.0, e1));
        st.dc = __dsub_rn(__dmul_rn(0.5, __dadd_rn(1.0, e2)), st.c);
    } else {
        st.c = __dmul_rn(0.5, erfc(-x));
        st.dc = __dsub_rn(__dmul_rn(0.5, erfc(-xb)), st.c);
    }
}

// ---------------------------------------------------------------------------
// Mixed per-dimension state (FP32 special functions, complementary form).
//
// With za, zb the clamped erf arguments of the interval, ea = erfc(|za|)/2 and
// eb = erfc(|zb|)/2 are the normal tail masses beyond |za|, |zb|.  Then
//   zb <= 0 (interval in the lower half):  Phi(a') = ea, dc = eb - ea
//   za >= 0 (upper half):                   1-Phi(b') = eb, dc = ea - eb
//   otherwise (straddles 0):                dc = 1 - ea - eb
// and the sample u = Phi(a') + w*dc is formed from whichever tail it is
// nearer to, so Phi^{-1} never sees a value close to 1 and the difference
// never cancels in the tails.
// ---------------------------------------------------------------------------
struct StateMixed {
    float ea, eb, dc;
    int mode;  // 0 lower, 1 upper, 2 straddle
};

__device__ __forceinline__ void mixed_from_z(StateMixed& st, float za, float zb)
{
    st.ea = 0.5f * erfcf(fabsf(za));
    st.eb = 0.5f * erfcf(fabsf(zb));
    if (zb <= 0.0f) {
        st.mode = 0;
        st.dc = st.eb - st.ea;
    } else if (za >= 0.0f) {
        st.mode = 1;
        st.dc = st.ea - st.eb;
    } else {
        st.mode = 2;
        st.dc = (1.0f - st.ea) - st.eb;
    }
}

// Phi^{-1} of a tail mass t <= 1/2 (lower tail), FP32 with an FP64 fallback
// for t < 1e-6.
__device__ __forceinline__ float ndtri_lower(float t)
{
    if (t < 1e-6f) return (float)(-kSqrt2 * erfcinv(2.0 * (double)t));
    return -1.41421356f * erfcinvf(2.0f * t);
}

// The tent-mapped point w in FP64, its complement, the clamp of the reference
// (u in [eps64, 1 - eps64]) and Phi^{-1}.
__device__ __forceinline__ float sample_mixed_w(const StateMixed& st, double w)
{
    const float wf = (float)w;
    const float wcf = (float)(1.0 - w);
    const float u_lo = fmaf(wf, st.dc, st.ea);   // Phi(a') + w dc          (modes 0, 2)
    const float u_hi = fmaf(wcf, st.dc, st.eb);  // 1 - (Phi(a') + w dc)    (modes 1, 2)
    const bool upper = (st.mode == 1) || (st.mode == 2 && u_hi < u_lo);
    float t = upper ? u_hi : u_lo;
    t = fmaxf(t, (float)kEps64);
    const float y = ndtri_lower(t);
    return upper ? -y : y;
}

__device__ __forceinline__ float sample_mixed(const StateMixed& st, double u)
{
    return sample_mixed_w(st, fabs(fma(u, 2.0, -1.0)));
}

__device__ __forceinline__ void phi_mixed(StateMixed& st, float s, float a, float b, float inv)
{
    float x = (a - s) * inv;
    float xb = (b - s) * inv;
    x = fmaxf(fminf(x, 9.0f), -9.0f);
    xb = fmaxf(fminf(xb, 9.0f), -9.0f);
    mixed_from_z(st, x, xb);
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wave_max_int(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ double wave_sum_f64(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide sum of pm * 2^pe as (S, E) with S * 2^E the sum.  Lanes with
// pm == 0 do not take part in the exponent maximum.  blockDim.x == NT.
template <int NT>
__device__ __forceinline__ void block_scaled_sum(double pm, int pe, double& S, int& E)
{
    __shared__ int s_e[NT / 64];
    __shared__ double s_v[NT / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int e = (pm != 0.0) ? pe : INT32_MIN;
    e = wave_max_int(e);
    if (lane == 0) s_e[wv] = e;
    __syncthreads();
    int Eb = s_e[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) Eb = max(Eb, s_e[w]);
    const int Euse = (Eb == INT32_MIN) ? 0 : Eb;
    double v = (pm != 0.0) ? ldexp(pm, pe - Euse) : pm;  // keeps NaN
    v = wave_sum_f64(v);
    if (lane == 0) s_v[wv] = v;
    __syncthreads();
    double tot = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) tot += s_v[w];
    S = tot;
    E = Euse;
    __syncthreads();
}

}  // namespace dbg
