// Device-side building blocks shared by the integration kernels: the fused
// Richtmyer / Sobol point map, the two per-dimension precisions (reference
// FP64 arithmetic and the FP32 mixed form), the scaled weight product and
// the block reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "dbg_internal.h"
#include "tailq.h"

namespace dbg {

constexpr double kSqrt2 = 1.4142135623730951;       // Float64(sqrt(2.0))
constexpr double kEps64 = 2.220446049250313e-16;    // eps(Float64)

// ---------------------------------------------------------------------------
// FP64 special functions behind call boundaries.  Inlined into the dimension
// loop, their polynomial constants are hoisted out of the loop into vector
// registers (two per FP64 constant), which exhausts the register file; as
// calls they keep the constants in scalar instructions local to each call.
// The FP32 ones of the mixed walk (tail mass, tail quantile) are inlined: their
// constants are single-register literals, and inlined the walk's scheduler can
// overlap their latency with the rest of the step (measured 7 % faster at
// d = 2500).
// ---------------------------------------------------------------------------
__device__ __attribute__((noinline)) double dbg_erfinv(double x);
__device__ __attribute__((noinline)) double dbg_erf(double x);
__device__ __attribute__((noinline)) double dbg_erfc(double x);
__device__ __attribute__((noinline)) double dbg_erfcinv(double x);

__device__ __attribute__((noinline)) double dbg_erfinv(double x) { return erfinv(x); }
__device__ __attribute__((noinline)) double dbg_erf(double x) { return erf(x); }
__device__ __attribute__((noinline)) double dbg_erfc(double x) { return erfc(x); }
__device__ __attribute__((noinline)) double dbg_erfcinv(double x) { return erfcinv(x); }
__device__ __forceinline__ float dbg_tail_quantile(float t, int k) { return tail_quantile_f32(t, k); }
// the same two behind call boundaries, for the sparse walk (a short loop body
// that is measured faster with calls)
__device__ __attribute__((noinline)) float dbg_tail_mass_call(float z, int& k) { return tail_mass_f32(z, k); }
__device__ __attribute__((noinline)) float dbg_tail_quantile_call(float t, int k) { return tail_quantile_f32(t, k); }

template <bool CALLS>
__device__ __forceinline__ float tail_mass_sel(float z, int& k)
{
    if constexpr (CALLS) return dbg_tail_mass_call(z, k);
    else return tail_mass_f32(z, k);
}

template <bool CALLS>
__device__ __forceinline__ float tail_quantile_sel(float t, int k)
{
    if constexpr (CALLS) return dbg_tail_quantile_call(t, k);
    else return dbg_tail_quantile(t, k);
}

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

// Per-dimension entries of the point tables (dimension g, shift k), read
// through the scalar cache.  The walks fetch them one dimension ahead of their
// use so that the load latency overlaps the previous dimension's work.
struct PointOp {
    uint64_t q64;  // Richtmyer, mixed: q in 64-bit fixed point
    uint32_t s32;  // Richtmyer, mixed: shift in 32-bit fixed point
    double q;      // Richtmyer, FP64: q
    double sf;     // Richtmyer, FP64: shift
    uint64_t su;   // Sobol: digital shift
};

template <bool MIXED, int GEN>
__device__ __forceinline__ PointOp point_op(const PointArgs& P, int g, int k)
{
    PointOp o{};
    const size_t gk = (size_t)g + (size_t)k * P.nd;
    if constexpr (GEN == kGenRichtmyer) {
        if constexpr (MIXED) {
            o.q64 = P.q64[g];
            o.s32 = P.shift32[gk];
        } else {
            o.q = P.q[g];
            o.sf = P.shifts_f[gk];
        }
    } else {
        o.su = P.shifts_u[gk];
    }
    return o;
}

// Uniform u in [0, 1] of one (point, dimension, shift) before the tent map:
// Richtmyer shift mod 1 (src/qmc_generators.jl:158-159) or Sobol digital shift
// (:197-198).  Computed in FP64 for both precisions.
template <int GEN>
__device__ __forceinline__ double point_u(const PointArgs& P, const PointOp& o, int g, int64_t j)
{
    if constexpr (GEN == kGenRichtmyer) {
        double u = __dadd_rn(o.sf, richtmyer_x(P.n_base + j, o.q));
        return u > 1.0 ? __dsub_rn(u, 1.0) : u;
    } else {
        const uint64_t mask = P.mask;
        const uint64_t shift = o.su & mask;
        const uint64_t x = (uint64_t)sobol_x(P.V + (size_t)g * 32, P.sobol_first + (uint64_t)j) & mask;
        return __dmul_rn((double)((x ^ shift) & mask), P.invscale);
    }
}

template <int GEN>
__device__ __forceinline__ double point_u(const PointArgs& P, int g, int k, int64_t j)
{
    return point_u<GEN>(P, point_op<false, GEN>(P, g, k), g, j);
}

// Sobol uniform of the FP64 path from the Sobol integer x = sobol_x(..)
__device__ __forceinline__ double sobol_u_from_x(const PointArgs& P, uint32_t x, uint64_t su)
{
    const uint64_t mask = P.mask;
    return __dmul_rn((double)((((uint64_t)x & mask) ^ (su & mask)) & mask), P.invscale);
}

// ---------------------------------------------------------------------------
// Sobol integers of a walk sub-block from LDS tables.  With idx = H 2^8 + l,
// the Gray code of idx is gray(l) ^ (bit 8 of idx) 2^7 in its low byte and
// gray(H) above it, so
//   sobol_x(V_g, idx) = low_g[gray(idx) & 255] ^ high_g(H),
//   low_g[e]  = XOR of V_g[k] over the set bits k < 8 of e,
//   high_g(H) = XOR of V_g[8 + b] over the set bits b of gray(H).
// The 256 points of a tile have consecutive indices, so H takes two values
// H0, H0 + 1 per tile.  The tables of SB dimensions (SB x 256 low entries,
// then SB x 2 high entries) are built by the workgroup before the walk.
// ---------------------------------------------------------------------------
struct SobolLane {
    uint32_t c;  // gray(idx) & 255, plus 256 (H - H0)
    __device__ __forceinline__ uint32_t lo() const { return c & 255u; }
    __device__ __forceinline__ uint32_t hi() const { return c >> 8; }
};

__device__ __forceinline__ SobolLane sobol_lane(const PointArgs& P, int64_t j0, int64_t j)
{
    const uint64_t idx = P.sobol_first + (uint64_t)j;
    const uint64_t H0 = (P.sobol_first + (uint64_t)j0) >> 8;
    SobolLane sl;
    sl.c = (uint32_t)((idx ^ (idx >> 1)) & 255u) | ((uint32_t)((idx >> 8) - H0) << 8);
    return sl;
}

// Sobol tables of a block of dimensions [gb, gb + cnt), cnt <= SB, for a tile
// whose first point index is j0 (NT >= 256 threads).  The Sobol integer of a
// lane's point in dimension gb + cc is the XOR of V[k] over the set bits k of
// its Gray code, split as bits 0-3, 4-7 (the low byte, sl.lo()) and 8-31 (the
// high part, common to the tile up to one carry, sl.hi()):
//   s_sob[cc * 32 + e]       XOR of V[k], k in 0..3, over the bits of e < 16
//   s_sob[cc * 32 + 16 + e]  XOR of V[4 + k], k in 0..3, over the bits of e
//   s_sob[SB * 32 + 2 cc + h] the high part for H0 + h (h = 0, 1)
// so that x = t[lo & 15] ^ t[16 + (lo >> 4)] ^ hi: the same XOR as one
// 256-entry table of the low byte, at a sixteenth of the table work.
constexpr int kSobTab = 32;  // words per dimension of the low tables

template <int SB>
constexpr int sobol_table_words()
{
    return SB * kSobTab + 2 * SB;
}

template <int SB, int NT>
__device__ __forceinline__ void sobol_tables(const PointArgs& P, int gb, int cnt, int64_t j0,
                                             uint32_t* __restrict__ s_sob, int tid)
{
    static_assert(NT >= 256 && SB * kSobTab <= 2 * NT && 2 * SB <= NT, "two low entries per thread");
#pragma unroll
    for (int rep = 0; rep < 2; ++rep) {
        const int ei = tid + rep * NT;  // entry: dimension cc, part (bits 0-3 or 4-7), index e
        const int cc = ei / kSobTab;
        if (ei < SB * kSobTab && cc < cnt) {
            const uint32_t e = (uint32_t)(ei & 15);
            const uint32_t* __restrict__ V = P.V + (size_t)(gb + cc) * 32 + 4 * ((ei >> 4) & 1);
            uint32_t x = 0u;
#pragma unroll
            for (int k = 0; k < 4; ++k) x ^= ((e >> k) & 1u) ? V[k] : 0u;
            s_sob[ei] = x;
        }
    }
    if (tid < 2 * cnt) {
        const int cc = tid >> 1;
        const uint64_t H = ((P.sobol_first + (uint64_t)j0) >> 8) + (uint64_t)(tid & 1);
        const uint32_t gh = (uint32_t)(H ^ (H >> 1));
        const uint32_t* __restrict__ V = P.V + (size_t)(gb + cc) * 32 + 8;
        uint32_t x = 0u;
#pragma unroll
        for (int b = 0; b < 24; ++b) x ^= ((gh >> b) & 1u) ? V[b] : 0u;
        s_sob[SB * kSobTab + tid] = x;
    }
}

// the lane's Sobol integer in dimension gb + cc from the block's tables
template <int SB>
__device__ __forceinline__ uint32_t sobol_from_tables(const uint32_t* __restrict__ s_sob, int cc, SobolLane sl)
{
    const uint32_t lo = sl.lo();
    return s_sob[cc * kSobTab + (lo & 15u)] ^ s_sob[cc * kSobTab + 16 + (lo >> 4)] ^
           s_sob[SB * kSobTab + 2 * cc + sl.hi()];
}

// The point-set tables enter each kernel as separate __restrict__ arguments:
// with that guarantee the compiler reads their per-dimension entries through
// the scalar cache (uniform addresses, no aliasing store) instead of as vector
// loads.  with_tables() puts them back into the PointArgs the helpers take.
#define DBG_POINT_PARAMS                                                                          \
    const double* __restrict__ pt_q, const double* __restrict__ pt_sf,                            \
        const uint64_t* __restrict__ pt_q64, const uint32_t* __restrict__ pt_s32,                 \
        const uint32_t* __restrict__ pt_V, const uint64_t* __restrict__ pt_su
#define DBG_POINT_ARGS(P) (P).q, (P).shifts_f, (P).q64, (P).shift32, (P).V, (P).shifts_u
#define DBG_WITH_TABLES(P) with_tables((P), pt_q, pt_sf, pt_q64, pt_s32, pt_V, pt_su)

__device__ __forceinline__ PointArgs with_tables(PointArgs P, DBG_POINT_PARAMS)
{
    P.q = pt_q;
    P.shifts_f = pt_sf;
    P.q64 = pt_q64;
    P.shift32 = pt_s32;
    P.V = pt_V;
    P.shifts_u = pt_su;
    return P;
}

// The same uniform as a 32-bit fixed-point fraction u32 * 2^-32 (mixed path).
// Richtmyer: frac((n_base + j) q) from q in 64-bit fixed point (exact modular
// product, within 2^-32 of the FP64 value for n_base + j < 2^21), plus the
// shift in 32-bit fixed point, wrapping mod 1.  Sobol: the masked XOR of
// the FP64 path, aligned to the top of the word.
template <int GEN>
__device__ __forceinline__ uint32_t point_u32(const PointArgs& P, const PointOp& o, int g, int64_t j)
{
    if constexpr (GEN == kGenRichtmyer) {
        const uint64_t Q = o.q64;
        const uint64_t nj = (uint64_t)(P.n_base + j);
        const uint32_t nlo = (uint32_t)nj, nhi = (uint32_t)(nj >> 32);
        const uint32_t qlo = (uint32_t)Q, qhi = (uint32_t)(Q >> 32);
        const uint32_t x = __umulhi(nlo, qlo) + nlo * qhi + nhi * qlo;
        return x + o.s32;
    } else {
        const uint64_t v = ((uint64_t)sobol_x(P.V + (size_t)g * 32, P.sobol_first + (uint64_t)j) ^ o.su) &
                           P.mask;
        return (uint32_t)(P.sobol_bits <= 32 ? (v << (32 - P.sobol_bits)) : (v >> (P.sobol_bits - 32)));
    }
}

template <int GEN>
__device__ __forceinline__ uint32_t point_u32(const PointArgs& P, int g, int k, int64_t j)
{
    return point_u32<GEN>(P, point_op<true, GEN>(P, g, k), g, j);
}

// Sobol fixed-point uniform of the mixed path from the Sobol integer x
__device__ __forceinline__ uint32_t sobol_u32_from_x(const PointArgs& P, uint32_t x, uint64_t su)
{
    const uint64_t v = ((uint64_t)x ^ su) & P.mask;
    return (uint32_t)(P.sobol_bits <= 32 ? (v << (32 - P.sobol_bits)) : (v >> (P.sobol_bits - 32)));
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
    // pm back to [1/2, 1) after products taken without normalization (the
    // exponent moves to pe; a power-of-two scale changes no product's rounding
    // while it stays normal)
    __device__ __forceinline__ void settle()
    {
        int e = 0;
        pm = frexp_(pm, &e);
        pe += (pm == T(0)) ? 0 : e;
    }
};

// ---------------------------------------------------------------------------
// Per-dimension arithmetic.  Three forms share one interface (struct Path):
//
//   reference FP64 (kPhiErf / kPhiErfc): the arithmetic of qmc_loop! in
//     Float64 -- c = (1 + erf(x))/2 or erfc(-x)/2 with x clamped to [-9, 9],
//     u = muladd(w, dc, c) clamped to [eps, 1 - eps], y = sqrt2 erfinv(2u - 1)
//     (src/qmc.jl:385-429, :579-593, src/qmc_generators.jl:156-164);
//   tail (kPhiTail; FP64, and the FP32 mixed path): no clamp; with za, zb the
//     erf arguments of the interval, ea = erfc(|za|)/2 and eb = erfc(|zb|)/2
//     are the normal tail masses beyond |za|, |zb|:
//        zb <= 0 (lower half):  Phi(a') = ea, dc = eb - ea
//        za >= 0 (upper half):  1 - Phi(b') = eb, dc = ea - eb
//        otherwise:             dc = 1 - ea - eb
//     and the sample's tail mass t is formed from whichever tail it is nearer
//     to (u_lo = ea + w dc or u_hi = eb + (1 - w) dc), y = -+sqrt2 erfcinv(2t).
//     Nothing cancels in the tails and no interval is cut at 9.
//   In FP32 the three masses carry a binary scale 2^-k, and both the masses
//   (tail_mass_f32) and Phi^{-1} of a scaled mass (tail_quantile_f32) are
//   computed in FP32 for any |z| (capped at 30) and any k: no FP64 and no
//   data-dependent branch in the walk's dimension step.
// ---------------------------------------------------------------------------
struct StateF64 {
    double c, dc;
};

template <typename T>
struct StateTail {
    T ea, eb, dc;
    int mode;  // 0 lower half, 1 upper half, 2 straddles 0
    int k;     // binary scale of ea, eb, dc (FP32 only)
};

// first-dimension values computed in FP64: {c1, dc1, ea1, eb1, d1, mode1}
__device__ __forceinline__ void tail_set64(double ea, double eb, double dc, int mode, StateTail<double>& st)
{
    st.ea = ea;
    st.eb = eb;
    st.dc = dc;
    st.mode = mode;
    st.k = 0;
}

__device__ __forceinline__ void tail_set64(double ea, double eb, double dc, int mode, StateTail<float>& st)
{
    const double mx = fmax(fmax(ea, eb), fabs(dc));
    int k = 0;
    if (mx > 0.0 && mx < 1e-30) {
        int e = 0;
        (void)frexp(mx, &e);
        k = -e;  // mx * 2^k in [0.5, 1)
    }
    st.ea = (float)ldexp(ea, k);
    st.eb = (float)ldexp(eb, k);
    st.dc = (float)ldexp(dc, k);
    st.mode = mode;
    st.k = k;
}

// erfc(|z|)/2 is 0 exactly for an infinite bound: skip the call
__device__ __forceinline__ void tail_masses64(double za, double zb, double& ea, double& eb, double& dc,
                                              int& mode)
{
    ea = 0.0;
    eb = 0.0;
    if (!isinf(za)) ea = 0.5 * dbg_erfc(fabs(za));
    if (!isinf(zb)) eb = 0.5 * dbg_erfc(fabs(zb));
    if (zb <= 0.0) {
        mode = 0;
        dc = eb - ea;
    } else if (za >= 0.0) {
        mode = 1;
        dc = ea - eb;
    } else {
        mode = 2;
        dc = (1.0 - ea) - eb;
    }
}

__device__ __forceinline__ void tail_from_z(StateTail<double>& st, double za, double zb)
{
    double ea, eb, dc;
    int mode;
    tail_masses64(za, zb, ea, eb, dc, mode);
    tail_set64(ea, eb, dc, mode, st);
}

// FP32 masses without FP64 or a branch: each bound's mass erfc(|z|)/2 comes as
// m 2^-k (tail_mass_f32); a one-sided interval takes the scale of its larger
// mass (the bound nearer 0), so that ea, eb and dc are formed with the
// smaller mass shifted into that scale; an interval that straddles 0 has k = 0.
template <bool CALLS = false>
__device__ __forceinline__ void tail_from_z(StateTail<float>& st, float za, float zb)
{
    int ka, kb;
    float ma = tail_mass_sel<CALLS>(fabsf(za), ka);
    float mb = tail_mass_sel<CALLS>(fabsf(zb), kb);
    if (isinf(za)) ma = 0.0f;
    if (isinf(zb)) mb = 0.0f;
    const bool lo = zb <= 0.0f, hi = !lo && za >= 0.0f;
    const int k = lo ? kb : (hi ? ka : 0);
    const float ea = ldexpf(ma, k - ka), eb = ldexpf(mb, k - kb);
    // the three widths are formed and one selected, so that the step has no
    // divergent branch here
    const float d_lo = eb - ea, d_hi = ea - eb, d_mid = (1.0f - ea) - eb;
    st.ea = ea;
    st.eb = eb;
    st.dc = lo ? d_lo : (hi ? d_hi : d_mid);
    st.mode = lo ? 0 : (hi ? 1 : 2);
    st.k = k;
}

// The same state when the lower bound is -inf (za = -inf): ea = 0, and the
// interval is either below 0 (mode 0, dc = eb) or straddles it (mode 2,
// dc = 1 - eb).  tail_from_z gives the same values for za = -inf (its lower
// mass is set to 0 and every select above reduces to these); here the lower
// bound's mass is never formed.
template <bool CALLS = false>
__device__ __forceinline__ void tail_from_zb(StateTail<float>& st, float zb)
{
    int kb;
    float mb = tail_mass_sel<CALLS>(fabsf(zb), kb);
    if (isinf(zb)) mb = 0.0f;
    const bool lo = zb <= 0.0f;
    const int k = lo ? kb : 0;
    const float eb = ldexpf(mb, k - kb);
    st.ea = 0.0f;
    st.eb = eb;
    st.dc = lo ? eb : 1.0f - eb;
    st.mode = lo ? 0 : 2;
    st.k = k;
}

__device__ __forceinline__ double tail_sample_w(const StateTail<double>& st, double w)
{
    const double u_lo = fma(w, st.dc, st.ea);
    const double u_hi = fma(1.0 - w, st.dc, st.eb);
    const bool upper = (st.mode == 1) || (st.mode == 2 && u_hi < u_lo);
    const double t = fmax(upper ? u_hi : u_lo, 1e-300);
    const double y = -kSqrt2 * dbg_erfcinv(2.0 * t);
    return upper ? -y : y;
}

// FP32 sample from the fixed-point point: w and 1 - w as FP32 values, the
// quantile of the nearer tail mass t * 2^-k in FP32 (log-domain, any k)
template <bool CALLS = false>
__device__ __forceinline__ float tail_sample_f(const StateTail<float>& st, float w, float wc)
{
    const float u_lo = fmaf(w, st.dc, st.ea);
    const float u_hi = fmaf(wc, st.dc, st.eb);
    const bool upper = (st.mode == 1) || (st.mode == 2 && u_hi < u_lo);
    const float t = fmaxf(upper ? u_hi : u_lo, 1.17549435e-38f);
    const float y = tail_quantile_sel<CALLS>(t, st.k);
    return upper ? -y : y;
}

__device__ __forceinline__ float tail_sample_w(const StateTail<float>& st, double w)
{
    return tail_sample_f<true>(st, (float)w, (float)(1.0 - w));
}

template <bool MIXED, int FORM>
struct Path;

// reference FP64 arithmetic (erf or erfc form)
template <int FORM>
struct Path<false, FORM> {
    using T = double;
    using State = StateF64;
    static __device__ __forceinline__ void init(State& st, Weight<T>& w, const double* f)
    {
        st.c = f[0];
        st.dc = f[1];
        w.init(f[1]);
    }
    // sample from the already tent-mapped point w (the SparseQMC module has
    // no tent map: src/qmc_sparse.jl:202-207)
    static __device__ __forceinline__ T sample_w(const State& st, double w)
    {
        double u = fma(w, st.dc, st.c);
        u = fmin(fmax(u, kEps64), 1.0 - kEps64);
        const double Yv = fma(2.0, u, -1.0);
        return __dmul_rn(kSqrt2, dbg_erfinv(Yv));
    }
    static __device__ __forceinline__ void phi(State& st, T s, T a, T b, T inv)
    {
        phi_z(st, __dmul_rn(__dsub_rn(a, s), inv), __dmul_rn(__dsub_rn(b, s), inv));
    }
    // the state of the interval with erf arguments x <= xb
    static __device__ __forceinline__ void phi_z(State& st, double x, double xb)
    {
        x = fmax(fmin(x, 9.0), -9.0);
        xb = fmax(fmin(xb, 9.0), -9.0);
        if constexpr (FORM == kPhiErf) {
            const double e1 = dbg_erf(x), e2 = dbg_erf(xb);
            st.c = __dmul_rn(0.5, __dadd_rn(1.0, e1));
            st.dc = __dsub_rn(__dmul_rn(0.5, __dadd_rn(1.0, e2)), st.c);
        } else {
            st.c = __dmul_rn(0.5, dbg_erfc(-x));
            st.dc = __dsub_rn(__dmul_rn(0.5, dbg_erfc(-xb)), st.c);
        }
    }
    static __device__ __forceinline__ void weigh(Weight<T>& w, const State& st) { w.mul(st.dc); }
    static __device__ __forceinline__ void settle(Weight<T>&) {}
};

// tail form in FP64
template <>
struct Path<false, kPhiTail> {
    using T = double;
    using State = StateTail<double>;
    static __device__ __forceinline__ void init(State& st, Weight<T>& w, const double* f)
    {
        tail_set64(f[2], f[3], f[4], (int)f[5], st);
        w.init(st.dc);
    }
    static __device__ __forceinline__ T sample_w(const State& st, double w) { return tail_sample_w(st, w); }
    static __device__ __forceinline__ void phi(State& st, T s, T a, T b, T inv)
    {
        tail_from_z(st, (a - s) * inv, (b - s) * inv);
    }
    static __device__ __forceinline__ void phi_z(State& st, double za, double zb) { tail_from_z(st, za, zb); }
    static __device__ __forceinline__ void weigh(Weight<T>& w, const State& st) { w.mul(st.dc); }
    static __device__ __forceinline__ void settle(Weight<T>&) {}
};

// the FP32 mixed path (tail form, any FORM; kPhiTailLower when every lower
// bound is -inf)
template <int FORM>
struct Path<true, FORM> {
    using T = float;
    using State = StateTail<float>;
    static __device__ __forceinline__ void init(State& st, Weight<T>& w, const double* f)
    {
        tail_set64(f[2], f[3], f[4], (int)f[5], st);
        w.init(st.dc);
        w.pe -= st.k;
    }
    static __device__ __forceinline__ T sample_w(const State& st, double w) { return tail_sample_w(st, w); }
    static __device__ __forceinline__ void phi(State& st, T s, T a, T b, T inv)
    {
        if constexpr (FORM == kPhiTailLower) tail_from_zb(st, (b - s) * inv);
        else tail_from_z(st, (a - s) * inv, (b - s) * inv);
    }
    // The lower-orthant form multiplies without normalizing: each factor
    // (m in [0.02, 0.71] of a tail mass m 2^-k, or 1 - m 2^-k >= 1/2, or 0)
    // keeps 16 products of a weight in [1/2, 1) above 2^-91, so the walks
    // settle the weight every 16 steps and at the end of each block, with
    // the products rounded exactly as with normalization at every step.
    static __device__ __forceinline__ void weigh(Weight<T>& w, const State& st)
    {
        if constexpr (FORM == kPhiTailLower) w.pm *= st.dc;
        else w.mul(st.dc);
        w.pe -= st.k;
    }
    static __device__ __forceinline__ void settle(Weight<T>& w)
    {
        if constexpr (FORM == kPhiTailLower) w.settle();
    }
};

// The first dimension's interval, analytic (src/qmc.jl:800-809), from its
// erf arguments arg_a = a_1 / (U_11 sqrt 2), arg_b likewise (unclamped):
// first = {c1, dc1} of the erf or erfc form, then {ea, eb, d, mode} of the
// tail form.
__device__ __forceinline__ void first_from_args(double arg_a, double arg_b, int phi_form, double* first)
{
    double c1, dc1;
    if (phi_form == kPhiErfc) {
        c1 = 0.5 * erfc(-arg_a);
        dc1 = 0.5 * (erfc(-arg_b) - erfc(-arg_a));
    } else {
        c1 = 0.5 * (1.0 + erf(arg_a));
        dc1 = 0.5 * (erf(arg_b) - erf(arg_a));
    }
    first[0] = c1;
    first[1] = dc1;
    // mixed form of the same interval (tail masses, complementary width)
    const double ea = 0.5 * erfc(fabs(arg_a));
    const double eb = 0.5 * erfc(fabs(arg_b));
    double d;
    int mode;
    if (arg_b <= 0.0) {
        mode = 0;
        d = eb - ea;
    } else if (arg_a >= 0.0) {
        mode = 1;
        d = ea - eb;
    } else {
        mode = 2;
        d = 0.5 * (erf(arg_b) - erf(arg_a));
    }
    first[2] = ea;
    first[3] = eb;
    first[4] = d;
    first[5] = (double)mode;
}

template <bool MIXED, int FORM>
__device__ __forceinline__ typename Path<MIXED, FORM>::T path_sample(const typename Path<MIXED, FORM>::State& st,
                                                                     double u)
{
    return Path<MIXED, FORM>::sample_w(st, fabs(fma(u, 2.0, -1.0)));
}

// y of dimension g for the lane's sample: point, tent map, inverse CDF.  The
// FP64 forms follow the reference's FP64 arithmetic; the mixed form takes
// the point in fixed point and the quantile in FP32.
template <bool MIXED, int GEN, int FORM>
__device__ __forceinline__ typename Path<MIXED, FORM>::T path_draw(const PointArgs& pts, const PointOp& op,
                                                                   const typename Path<MIXED, FORM>::State& st,
                                                                   int g, int64_t j)
{
    if constexpr (MIXED) {
        float w, wc;
        tent_w(point_u32<GEN>(pts, op, g, j), w, wc);
        return tail_sample_f(st, w, wc);
    } else {
        return path_sample<MIXED, FORM>(st, point_u<GEN>(pts, op, g, j));
    }
}

// the same draw for a Sobol point given its Sobol integer x
template <bool MIXED, int FORM>
__device__ __forceinline__ typename Path<MIXED, FORM>::T path_draw_sobol(const PointArgs& pts, const PointOp& op,
                                                                         const typename Path<MIXED, FORM>::State& st,
                                                                         uint32_t x)
{
    if constexpr (MIXED) {
        float w, wc;
        tent_w(sobol_u32_from_x(pts, x, op.su), w, wc);
        return tail_sample_f(st, w, wc);
    } else {
        return path_sample<MIXED, FORM>(st, sobol_u_from_x(pts, x, op.su));
    }
}

// ---------------------------------------------------------------------------
// Walk of one block of up to SB sampled dimensions [gb, gb + cnt) for the
// lane's sample (lane t of a tile of TS samples).
//
// On entry s_x[t * SXS + c] holds the conditional mean of target gb + 1 + c
// accumulated from every dimension before gb.  For each dimension g the lane
// draws the point, forms y_g, stores it (LDS row s_y[g - gb], global
// workspace row Ys[g]), completes the mean of target g + 1 with the block's
// own y values in ascending order (the in-block dot of src/qmc.jl:390-403,
// y held in registers: the loop is unrolled), and updates its interval and
// weight (:405-429).  Ucol is the factor in column-major order
// (Ucol[j * ldc + l] = U[l, j]), read through the scalar cache since every
// lane needs the same column.
// ---------------------------------------------------------------------------
// ys: row stride of the y workspace Ys; srec (optional): the conditional
// means s_{g+1} are also recorded there (same stride), for the adjoint;
// s_sob (optional, Sobol): the block's Sobol tables from sobol_tables, with
// the lane's SobolLane.  Each step's scalar operands are fetched one step
// ahead (StepOps).
//
// Scalar operands of one walk step (sampled dimension g of the block at gb):
// the point-table entries, column g + 1 of the factor from row gb to g
// (the in-block dot), and the bounds of target g + 1.
template <bool MIXED, int GEN, int SB, typename T>
struct StepOps {
    PointOp pt;
    T u[SB];
    T a, b, inv;

    template <int CC>
    __device__ __forceinline__ void fetch(const PointArgs& pts, int gb, int k, const T* __restrict__ Ucol,
                                          int64_t ldc, const T* __restrict__ ak, const T* __restrict__ bk,
                                          const T* __restrict__ invk)
    {
        const int g = gb + CC;
        pt = point_op<MIXED, GEN>(pts, g, k);
        const T* __restrict__ ucol = Ucol + (size_t)(g + 1) * ldc + gb;
#pragma unroll
        for (int l = 0; l <= CC; ++l) u[l] = ucol[l];
        a = ak[g + 1];
        b = bk[g + 1];
        inv = invk[g + 1];
    }
};

template <bool MIXED, int GEN, int FORM, int TS, int SXS, int SB, int CC, bool YST = true, bool SX = false,
          typename T, typename State>
__device__ __forceinline__ void walk_steps(const PointArgs& pts, int gb, int cnt, int k, int64_t j, int t,
                                           State& st, Weight<T>& w, T (&yv)[SB], StepOps<MIXED, GEN, SB, T>& cur,
                                           const T* __restrict__ s_x, T* __restrict__ s_y, T* __restrict__ Ys,
                                           int64_t ys, const T* __restrict__ Ucol, int64_t ldc,
                                           const T* __restrict__ ak, const T* __restrict__ bk,
                                           const T* __restrict__ invk, T* __restrict__ srec,
                                           const uint32_t* __restrict__ s_sob, SobolLane sl)
{
    if constexpr (CC < SB) {
        if (CC < cnt) {  // uniform
            using PT = Path<MIXED, FORM>;
            StepOps<MIXED, GEN, SB, T> nxt;
            if constexpr (CC + 1 < SB) {
                // the next step's operands, in flight while this step computes
                if (CC + 1 < cnt) nxt.template fetch<CC + 1>(pts, gb, k, Ucol, ldc, ak, bk, invk);
            }
            const int g = gb + CC;
            T y;
            if (GEN == kGenSobol && s_sob != nullptr)
                y = path_draw_sobol<MIXED, FORM>(pts, cur.pt, st,
                                                 sobol_from_tables<SB>(s_sob, CC, sl));
            else
                y = path_draw<MIXED, GEN, FORM>(pts, cur.pt, st, g, j);
            yv[CC] = y;
            s_y[CC * TS + t] = y;
            if constexpr (YST) Ys[(size_t)g * ys + t] = y;
            T acc = s_x[t * SXS + CC];
#pragma unroll
            for (int l = 0; l <= CC; ++l) acc = fma(yv[l], cur.u[l], acc);
            if (srec) srec[(size_t)g * ys + t] = acc;
            // SX: the conditional mean back into the hand-off slot it came from
            // (read only by this thread), for the caller to record
            if constexpr (SX) const_cast<T*>(s_x)[t * SXS + CC] = acc;
            PT::phi(st, acc, cur.a, cur.b, cur.inv);
            PT::weigh(w, st);
            if constexpr (CC % 16 == 15) PT::settle(w);
            if constexpr (CC + 1 < SB)
                walk_steps<MIXED, GEN, FORM, TS, SXS, SB, CC + 1, YST, SX>(pts, gb, cnt, k, j, t, st, w, yv, nxt, s_x,
                                                                       s_y, Ys, ys, Ucol, ldc, ak, bk, invk, srec,
                                                                       s_sob, sl);
        }
    }
}

// The walk of one sub-block; with YST = false the y values are not stored
// to the workspace row by row but left in yv for the caller (which stores
// the whole sub-block at once in its own layout).
template <bool MIXED, int GEN, int FORM, int TS, int SXS, int SB, bool YST, bool SX = false, typename T,
          typename State>
__device__ __forceinline__ void walk_block_yv(T (&yv)[SB], const PointArgs& pts, int gb, int cnt, int k,
                                              int64_t j, int t, State& st, Weight<T>& w,
                                              const T* __restrict__ s_x, T* __restrict__ s_y,
                                              T* __restrict__ Ys, int64_t ys, const T* __restrict__ Ucol,
                                              int64_t ldc, const T* __restrict__ ak, const T* __restrict__ bk,
                                              const T* __restrict__ invk, T* __restrict__ srec = nullptr,
                                              const uint32_t* __restrict__ s_sob = nullptr, SobolLane sl = {})
{
    StepOps<MIXED, GEN, SB, T> first;
    if (cnt > 0) first.template fetch<0>(pts, gb, k, Ucol, ldc, ak, bk, invk);
    walk_steps<MIXED, GEN, FORM, TS, SXS, SB, 0, YST, SX>(pts, gb, cnt, k, j, t, st, w, yv, first, s_x, s_y, Ys, ys,
                                                      Ucol, ldc, ak, bk, invk, srec, s_sob, sl);
    if constexpr (SB % 16 != 0) Path<MIXED, FORM>::settle(w);
    else if (cnt % 16 != 0) Path<MIXED, FORM>::settle(w);  // uniform: a short last block
}

template <bool MIXED, int GEN, int FORM, int TS, int SXS, int SB, typename T, typename State>
__device__ __forceinline__ void walk_block(const PointArgs& pts, int gb, int cnt, int k, int64_t j,
                                           int t, State& st, Weight<T>& w, const T* __restrict__ s_x,
                                           T* __restrict__ s_y, T* __restrict__ Ys, int64_t ys,
                                           const T* __restrict__ Ucol, int64_t ldc,
                                           const T* __restrict__ ak, const T* __restrict__ bk,
                                           const T* __restrict__ invk, T* __restrict__ srec = nullptr,
                                           const uint32_t* __restrict__ s_sob = nullptr, SobolLane sl = {})
{
    T yv[SB];
    walk_block_yv<MIXED, GEN, FORM, TS, SXS, SB, true>(yv, pts, gb, cnt, k, j, t, st, w, s_x, s_y, Ys, ys, Ucol,
                                                       ldc, ak, bk, invk, srec, s_sob, sl);
}

// The same walk for a kernel whose lane computes its own sample's conditional
// means in registers (the vector-ALU contraction of the batched kernel): the
// hand-off sx[c] (target gb + 1 + c) is a register array and y goes only to
// the global workspace Ys (row g, stride ys), so the walk needs no LDS.  Same
// arithmetic and operation order as walk_steps.
template <bool MIXED, int GEN, int FORM, int SB, int CC, bool YST, typename T, typename State>
__device__ __forceinline__ void walk_steps_reg(const PointArgs& pts, int gb, int cnt, int k, int64_t j, int t,
                                               State& st, Weight<T>& w, T (&yv)[SB], const T (&sx)[SB],
                                               StepOps<MIXED, GEN, SB, T>& cur, T* __restrict__ Ys, int64_t ys,
                                               const T* __restrict__ Ucol, int64_t ldc, const T* __restrict__ ak,
                                               const T* __restrict__ bk, const T* __restrict__ invk)
{
    if constexpr (CC < SB) {
        if (CC < cnt) {  // uniform
            using PT = Path<MIXED, FORM>;
            StepOps<MIXED, GEN, SB, T> nxt;
            if constexpr (CC + 1 < SB) {
                if (CC + 1 < cnt) nxt.template fetch<CC + 1>(pts, gb, k, Ucol, ldc, ak, bk, invk);
            }
            const int g = gb + CC;
            const T y = path_draw<MIXED, GEN, FORM>(pts, cur.pt, st, g, j);
            yv[CC] = y;
            if constexpr (YST) Ys[(size_t)g * ys + t] = y;
            T acc = sx[CC];
#pragma unroll
            for (int l = 0; l <= CC; ++l) acc = fma(yv[l], cur.u[l], acc);
            PT::phi(st, acc, cur.a, cur.b, cur.inv);
            PT::weigh(w, st);
            if constexpr (CC % 16 == 15) PT::settle(w);
            if constexpr (CC + 1 < SB)
                walk_steps_reg<MIXED, GEN, FORM, SB, CC + 1, YST>(pts, gb, cnt, k, j, t, st, w, yv, sx, nxt, Ys, ys,
                                                                  Ucol, ldc, ak, bk, invk);
        }
    }
}

// YST = false leaves the block's y in yv for the caller instead of storing
// them to Ys.
template <bool MIXED, int GEN, int FORM, int SB, bool YST = true, typename T, typename State>
__device__ __forceinline__ void walk_block_reg(const PointArgs& pts, int gb, int cnt, int k, int64_t j, int t,
                                               State& st, Weight<T>& w, const T (&sx)[SB], T (&yv)[SB],
                                               T* __restrict__ Ys, int64_t ys, const T* __restrict__ Ucol,
                                               int64_t ldc, const T* __restrict__ ak, const T* __restrict__ bk,
                                               const T* __restrict__ invk)
{
    StepOps<MIXED, GEN, SB, T> first;
    if (cnt > 0) first.template fetch<0>(pts, gb, k, Ucol, ldc, ak, bk, invk);
    walk_steps_reg<MIXED, GEN, FORM, SB, 0, YST>(pts, gb, cnt, k, j, t, st, w, yv, sx, first, Ys, ys, Ucol, ldc, ak,
                                                 bk, invk);
    if constexpr (SB % 16 != 0) Path<MIXED, FORM>::settle(w);
    else if (cnt % 16 != 0) Path<MIXED, FORM>::settle(w);  // uniform: a short last block
}

// ---------------------------------------------------------------------------
// Reverse sweep (adjoint of the tail-form estimate; no reference counterpart).
//
// For one sample, with alpha_r = (a_r - s_r)/U_rr, beta_r = (b_r - s_r)/U_rr,
// dc_r = Phi(beta_r) - Phi(alpha_r), y_r = Phi^{-1}(Phi(alpha_r) + w_r dc_r)
// and s_r = sum_{l<r} U[l, r] y_l, the reverse step of index r (given the
// complete ybar_r = sum_{r'>r} U[r, r'] sbar_r') is
//   dcbar = wt / dc_r + ybar_r w_r / phi(y_r),   cbar = ybar_r / phi(y_r)
//   albar = (cbar - dcbar) phi(alpha_r),         bebar = dcbar phi(beta_r)
//   abar_r = albar / U_rr, bbar_r = bebar / U_rr,
//   Ubar_rr = -(albar alpha_r + bebar beta_r) / U_rr,
//   sbar_r = -(albar + bebar) / U_rr,
// with wt the sample's normalised weight p / sum(p), so that the sums over
// samples are d log P directly.  The masses and densities carry the state's
// binary scale 2^k (FP32), which cancels in every product above.  The terms
// of y_r are absent for r = d - 1, and s_0 = 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum_f64(double v);
__device__ __forceinline__ float wave_sum_f32(float v);

template <typename T>
__device__ __forceinline__ T scaled_pdf(T x, int k)
{
    if (isinf(x)) return T(0);
    if constexpr (std::is_same<T, float>::value)  // 2^(k - x^2 log2(e) / 2) by v_exp_f32 (1 ulp)
        return __builtin_amdgcn_exp2f(fmaf(-0.721347520444481703f * x, x, (float)k)) * 0.398942280401432678f;
    else
        return exp(-0.5 * x * x + k * 0.69314718055994530942) * 0.39894228040143267794;
}

struct RevOut {
    double ga, gb, gd;
};

// LOWER (mixed only): every lower bound is -inf, so alpha_r = -inf, phi(alpha_r)
// = 0 and the alpha terms vanish; the lower bound's mass and density are not
// formed (the same values as the general step with a_r = -inf).
template <bool MIXED, int GEN, bool LOWER = false, typename T = typename std::conditional<MIXED, float, double>::type>
__device__ __forceinline__ T rev_step(const PointArgs& pts, const PointOp& op, int n, int r, int64_t j, T ybar,
                                      T wt, T s_r, T y_r, T a_r, T b_r, T inv_r, RevOut& o)
{
    static_assert(MIXED || !LOWER, "the lower-orthant step is a form of the mixed path");
    const T zb = (b_r - s_r) * inv_r;
    const T za = LOWER ? T(0) : (a_r - s_r) * inv_r;
    typename Path<MIXED, kPhiTail>::State st;
    if constexpr (LOWER) tail_from_zb(st, zb);
    else tail_from_z(st, za, zb);
    const T sq2 = T(1.4142135623730951);
    const T al = za * sq2, be = zb * sq2;
    T dcb = (wt > T(0) && st.dc > T(0)) ? wt / st.dc : T(0);
    T cb = T(0);
    if (r <= n - 2) {
        T w;
        if constexpr (MIXED) {
            float wf, wc;
            tent_w(point_u32<GEN>(pts, op, r, j), wf, wc);
            w = wf;
        } else {
            w = fabs(fma(point_u<GEN>(pts, op, r, j), 2.0, -1.0));
        }
        const T ph = scaled_pdf(y_r, st.k);
        const T iy = ph > T(0) ? ybar / ph : T(0);
        cb = iy;
        dcb += iy * w;
    }
    const T albar = LOWER ? T(0) : (cb - dcb) * scaled_pdf(al, st.k);
    const T bebar = dcb * scaled_pdf(be, st.k);
    const T invU = inv_r * sq2;
    o.ga = (double)(albar * invU);
    o.gb = (double)(bebar * invU);
    const T ta = (LOWER || isinf(al)) ? T(0) : albar * al;
    const T tb = isinf(be) ? T(0) : bebar * be;
    o.gd = (double)(-(ta + tb) * invU);
    return -(albar + bebar) * invU;
}

template <bool MIXED, int GEN, bool LOWER = false, typename T = typename std::conditional<MIXED, float, double>::type>
__device__ __forceinline__ T rev_step(const PointArgs& pts, int n, int r, int k, int64_t j, T ybar, T wt,
                                      T s_r, T y_r, T a_r, T b_r, T inv_r, RevOut& o)
{
    const PointOp op = (r <= n - 2) ? point_op<MIXED, GEN>(pts, r, k) : PointOp{};
    return rev_step<MIXED, GEN, LOWER>(pts, op, n, r, j, ybar, wt, s_r, y_r, a_r, b_r, inv_r, o);
}

// Scalar operands of one reverse step (flipped index g = gb + CC, index
// r = n - 1 - g): the point-table entries of r, the bounds of r and column
// g + 1 of the flipped factor from row gb to g.
template <bool MIXED, int GEN, int SB, typename T>
struct RevOps {
    PointOp pt;
    T u[SB];
    T a, b, inv;

    template <int CC>
    __device__ __forceinline__ void fetch(const PointArgs& pts, int n, int gb, int k, const T* __restrict__ Uf,
                                          int64_t ldc, const T* __restrict__ ak, const T* __restrict__ bk,
                                          const T* __restrict__ invk)
    {
        const int g = gb + CC;
        const int r = n - 1 - g;
        pt = (r <= n - 2) ? point_op<MIXED, GEN>(pts, r, k) : PointOp{};
        const T* __restrict__ ucol = Uf + (size_t)(g + 1) * ldc + gb;
#pragma unroll
        for (int l = 0; l <= CC; ++l) u[l] = ucol[l];
        a = ak[r];
        b = bk[r];
        inv = invk[r];
    }
};

// unsc: the factor that undoes a per-sample scale of wt (and so of every
// sbar and ybar of the sample) where values leave the sample: the recorded
// sbar and the per-index sums (1 without such a scale).  YST = false leaves
// the sub-block's sbar in sv for the caller instead of storing them to Ys;
// SREC = false leaves the sbar records to the caller as well.  LOWER: the
// lower-orthant step (rev_step).
template <bool MIXED, int GEN, int TS, int SXS, int SB, int CC, bool YST = true, bool SREC = true,
          bool LOWER = false, typename T>
__device__ __forceinline__ void rev_steps(const PointArgs& pts, int n, int gb, int cnt, int k, int64_t j, int t,
                                          T& ybar, T wt, const T* __restrict__ s_x, T* __restrict__ s_y,
                                          T* __restrict__ Ys, const T* __restrict__ Uf, int64_t ldc,
                                          const T* __restrict__ ak, const T* __restrict__ bk,
                                          const T* __restrict__ invk, T* __restrict__ srec,
                                          const T* __restrict__ yrec, int64_t rs, double* __restrict__ red,
                                          T (&sv)[SB], T& s_nx, T& y_nx, RevOps<MIXED, GEN, SB, T> (&ops)[2],
                                          bool a_inf, T unsc)
{
    if constexpr (CC < SB) {
        if (CC < cnt) {  // uniform
            const int g = gb + CC;
            const int r = n - 1 - g;
            const T s_r = s_nx;
            const T y_r = y_nx;
            const RevOps<MIXED, GEN, SB, T>& cur = ops[CC & 1];
            if constexpr (CC + 1 < SB) {
                if (CC + 1 < cnt) {
                    // records stream from HBM once: non-temporal loads
                    s_nx = __builtin_nontemporal_load(&srec[(size_t)(r - 2) * rs + t]);
                    y_nx = __builtin_nontemporal_load(&yrec[(size_t)(r - 1) * rs + t]);  // r - 1 <= n - 2
                    ops[(CC + 1) & 1].template fetch<CC + 1>(pts, n, gb, k, Uf, ldc, ak, bk, invk);
                }
            }
            RevOut o;
            const T sbar =
                rev_step<MIXED, GEN, LOWER>(pts, cur.pt, n, r, j, ybar, wt, s_r, y_r, cur.a, cur.b, cur.inv, o);
            sv[CC] = sbar;
            s_y[CC * TS + t] = sbar;
            if constexpr (YST) Ys[(size_t)g * TS + t] = sbar;
            if constexpr (SREC) srec[(size_t)(r - 1) * rs + t] = sbar * unsc;
            // per-index sums over the wavefront: FP32 partials of 64 FP32
            // terms on the mixed path, FP64 on the FP64 path
            double a0 = 0.0, b0, d0;
            if constexpr (MIXED) {
                if (!LOWER && !a_inf) a0 = wave_sum_f32((float)o.ga * unsc);  // uniform; 0 when a' = -inf throughout
                b0 = wave_sum_f32((float)o.gb * unsc);
                d0 = wave_sum_f32((float)o.gd * unsc);
            } else {
                if (!a_inf) a0 = wave_sum_f64(o.ga * unsc);
                b0 = wave_sum_f64(o.gb * unsc);
                d0 = wave_sum_f64(o.gd * unsc);
            }
            if ((t & 63) == 0) {
                double* rr = red + ((t >> 6) * SB + CC) * 3;
                rr[0] = a0;
                rr[1] = b0;
                rr[2] = d0;
            }
            T acc = s_x[t * SXS + CC];
#pragma unroll
            for (int l = 0; l <= CC; ++l) acc = fma(sv[l], cur.u[l], acc);
            ybar = acc;
            rev_steps<MIXED, GEN, TS, SXS, SB, CC + 1, YST, SREC, LOWER>(pts, n, gb, cnt, k, j, t, ybar, wt, s_x,
                                                                         s_y, Ys, Uf, ldc, ak, bk, invk, srec, yrec,
                                                                         rs, red, sv, s_nx, y_nx, ops, a_inf, unsc);
        }
    }
}

// Reverse walk of flipped indices [gb, gb + cnt) (flipped index g is r =
// n - 1 - g).  On entry s_x[t * SXS + c] holds the part of ybar for flipped
// index gb + c accumulated from every flipped index before gb; Uf is the
// flipped factor Uf[g1, g2] = U[n-1-g2, n-1-g1] (column-major, ldc), so the
// in-block part is the same ascending dot as in the forward walk.  sbar goes
// to the flipped workspace (s_y, Ys) for the contraction and to the record
// srec[(r - 1) * rs + t] in place of s_r for the gradient of U.  Per index,
// the three sums over the workgroup's samples are left in red[wave][cc][3].
// The records of a step are read one step ahead (they stream from HBM): on
// entry s_nx, y_nx hold those of the block's first step (record_first_step);
// step cc overwrites record row r - 1 only after reading it, and the next
// step reads row r - 2.  The scalar operands are fetched one step ahead as
// well.
template <bool MIXED, int GEN, int TS, int SXS, int SB, bool YST, bool SREC = true, bool LOWER = false, typename T>
__device__ __forceinline__ void walk_block_rev_sv(T (&sv)[SB], const PointArgs& pts, int n, int gb, int cnt, int k,
                                                  int64_t j, int t, T& ybar, T wt, const T* __restrict__ s_x,
                                                  T* __restrict__ s_y, T* __restrict__ Ys,
                                                  const T* __restrict__ Uf, int64_t ldc,
                                                  const T* __restrict__ ak, const T* __restrict__ bk,
                                                  const T* __restrict__ invk, T* __restrict__ srec,
                                                  const T* __restrict__ yrec, int64_t rs,
                                                  double* __restrict__ red, bool a_inf, T unsc, T s_nx, T y_nx)
{
    RevOps<MIXED, GEN, SB, T> ops[2];
    if (cnt > 0) ops[0].template fetch<0>(pts, n, gb, k, Uf, ldc, ak, bk, invk);
    rev_steps<MIXED, GEN, TS, SXS, SB, 0, YST, SREC, LOWER>(pts, n, gb, cnt, k, j, t, ybar, wt, s_x, s_y, Ys, Uf, ldc,
                                                            ak, bk, invk, srec, yrec, rs, red, sv, s_nx, y_nx, ops,
                                                            a_inf, unsc);
}

// The records of the first step of the reverse block at flipped index gb
// (r0 = n - 1 - gb >= 1): s_r0 from row r0 - 1, y_r0 from row r0 (none for
// r0 = n - 1).
template <typename T>
__device__ __forceinline__ void record_first_step(int n, int gb, int t, const T* __restrict__ srec,
                                                  const T* __restrict__ yrec, int64_t rs, T& s_nx, T& y_nx)
{
    const int r0 = n - 1 - gb;
    s_nx = __builtin_nontemporal_load(&srec[(size_t)(r0 - 1) * rs + t]);
    y_nx = (r0 <= n - 2) ? __builtin_nontemporal_load(&yrec[(size_t)r0 * rs + t]) : T(0);
}

template <bool MIXED, int GEN, int TS, int SXS, int SB, typename T>
__device__ __forceinline__ void walk_block_rev(const PointArgs& pts, int n, int gb, int cnt, int k, int64_t j,
                                               int t, T& ybar, T wt, const T* __restrict__ s_x,
                                               T* __restrict__ s_y, T* __restrict__ Ys,
                                               const T* __restrict__ Uf, int64_t ldc,
                                               const T* __restrict__ ak, const T* __restrict__ bk,
                                               const T* __restrict__ invk, T* __restrict__ srec,
                                               const T* __restrict__ yrec, int64_t rs,
                                               double* __restrict__ red, bool a_inf)
{
    T sv[SB];
    T s_nx = T(0), y_nx = T(0);
    if (cnt > 0) record_first_step(n, gb, t, srec, yrec, rs, s_nx, y_nx);
    walk_block_rev_sv<MIXED, GEN, TS, SXS, SB, true>(sv, pts, n, gb, cnt, k, j, t, ybar, wt, s_x, s_y, Ys, Uf, ldc,
                                                      ak, bk, invk, srec, yrec, rs, red, a_inf, T(1), s_nx, y_nx);
}

// ---------------------------------------------------------------------------
// Matrix-core instruction wrappers (v_mfma_*_16x16x4)
// ---------------------------------------------------------------------------
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
// Result layout of the 16x16x4 instructions (checked on the device): lane l
// holds column l % 16; element r of its accumulator is row 4 * (l / 16) + r
// for f32 and row (l / 16) + 4 * r for f64.
template <>
struct Mfma<float> {
    using V = floatx4;
    static __device__ __forceinline__ int row(int lq, int r) { return 4 * lq + r; }
    static __device__ __forceinline__ V mac(float a, float b, V c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mfma<double> {
    using V = doublex4;
    static __device__ __forceinline__ int row(int lq, int r) { return lq + 4 * r; }
    static __device__ __forceinline__ V mac(double a, double b, V c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};


// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wave_max_int(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// Sum over the 64 lanes with DPP moves (vector ALU only, no LDS round trips):
// row_shr 1, 2, 4, 8 leave each 16-lane row's sum in its lane 15 (lanes
// without a source read 0), row_bcast 15 / 31 carry rows 0 and 0-1 into rows
// 1, 3 and 2-3, so lane 63 ends with the total, which is then broadcast.
template <int CTRL, int ROW_MASK, bool BOUND_ZERO>
__device__ __forceinline__ float dpp_f32(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROW_MASK,
                                                                 0xF, BOUND_ZERO));
}

__device__ __forceinline__ float wave_sum_f32(float v)
{
    v += dpp_f32<0x111, 0xF, true>(v);  // row_shr:1
    v += dpp_f32<0x112, 0xF, true>(v);  // row_shr:2
    v += dpp_f32<0x114, 0xF, true>(v);  // row_shr:4
    v += dpp_f32<0x118, 0xF, true>(v);  // row_shr:8
    v += dpp_f32<0x142, 0xA, false>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_f32<0x143, 0xC, false>(v);  // row_bcast:31 into rows 2 and 3
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
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
