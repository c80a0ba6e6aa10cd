// FP64 re-evaluation of the heaviest tiles of a mixed-precision round (no
// reference counterpart).
//
// A round's estimate of shift k is the mean of its per-sample weights, and
// in high dimension those weights span hundreds of orders of magnitude: at
// d = 2500 (factor model, lower orthant) the standard deviation of log w over
// the samples is about 100, so a handful of samples carries all but 1e-8 of
// the shift's sum.  The FP32 walk computes each weight to about 1e-3 relative
// there (rounding of the conditional means and of z, coherent over 2500
// dimensions), which the estimate inherits from its few dominant samples.
//
// After the mixed kernel has written its per-tile partial sums, one
// workgroup per shift picks the K tiles with the largest sums (ties: lower
// tile index first) and records the fraction of the shift's sum that the
// other tiles hold.  The FP64 integration kernel then runs exactly those tiles
// (an FP64 tile covers 1 / ratio of a mixed tile) with the FP64 factor, FP64
// points and FP64 special functions -- the arithmetic of the plain FP64 path
// -- and the merge replaces each picked tile's partial sum by the sum of its
// FP64 tiles.  The estimate then differs from the FP64 path's only through
// the unpicked tiles, whose share of the sum is reported per shift.
#include "dbg_device.h"

namespace dbg {

namespace {

// log2 of a tile partial S * 2^E; -inf for an empty or non-finite tile
__device__ __forceinline__ double tile_key(double S, int E)
{
    return (S > 0.0 && isfinite(S)) ? log2(S) + (double)E : -INFINITY;
}

// order of the selection: larger key first, then smaller index
__device__ __forceinline__ bool before(double ka, int64_t ia, double kb, int64_t ib)
{
    return ka > kb || (ka == kb && ia < ib);
}

}  // namespace

__global__ __launch_bounds__(256) void refine_select_kernel(const double* __restrict__ part_m,
                                                            const int* __restrict__ part_e, int64_t tpr, int K,
                                                            int ratio, int64_t tpr64, int64_t* __restrict__ sel,
                                                            int64_t* __restrict__ list64, double* __restrict__ frac)
{
    __shared__ double s_k[4];
    __shared__ int64_t s_i[4];
    __shared__ int s_e[4];
    __shared__ double s_s[4];
    const int k = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const double* pm = part_m + (int64_t)k * tpr;
    const int* pe = part_e + (int64_t)k * tpr;
    double lk = INFINITY;  // the last pick: the next one comes after it in the order
    int64_t li = -1;
    int picked = 0;
    for (int r = 0; r < K; ++r) {
        double bk = -INFINITY;
        int64_t bi = INT64_MAX;
        for (int64_t t = tid; t < tpr; t += 256) {
            const double key = tile_key(pm[t], pe[t]);
            if (key == -INFINITY || !before(lk, li, key, t)) continue;
            if (before(key, t, bk, bi)) {
                bk = key;
                bi = t;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ok = __shfl_xor(bk, off, 64);
            const int64_t oi = __shfl_xor(bi, off, 64);
            if (before(ok, oi, bk, bi)) {
                bk = ok;
                bi = oi;
            }
        }
        if (lane == 0) {
            s_k[wv] = bk;
            s_i[wv] = bi;
        }
        __syncthreads();
        bk = s_k[0];
        bi = s_i[0];
        for (int w = 1; w < 4; ++w)
            if (before(s_k[w], s_i[w], bk, bi)) {
                bk = s_k[w];
                bi = s_i[w];
            }
        __syncthreads();
        if (bk == -INFINITY) break;  // uniform: no nonzero tile left
        if (tid == 0) sel[(int64_t)k * K + r] = (int64_t)k * tpr + bi;
        if (tid < ratio) {
            const int64_t t64 = (int64_t)ratio * bi + tid;
            list64[((int64_t)k * K + r) * ratio + tid] = t64 < tpr64 ? (int64_t)k * tpr64 + t64 : -1;
        }
        lk = bk;
        li = bi;
        picked = r + 1;
    }
    for (int r = picked + tid; r < K; r += 256) sel[(int64_t)k * K + r] = -1;
    for (int64_t e = (int64_t)picked * ratio + tid; e < (int64_t)K * ratio; e += 256)
        list64[(int64_t)k * K * ratio + e] = -1;
    // share of the shift's sum outside the picks: every tile after the last pick
    int E = INT32_MIN;
    for (int64_t t = tid; t < tpr; t += 256)
        if (tile_key(pm[t], pe[t]) != -INFINITY) E = max(E, pe[t]);
    E = wave_max_int(E);
    if (lane == 0) s_e[wv] = E;
    __syncthreads();
    E = max(max(s_e[0], s_e[1]), max(s_e[2], s_e[3]));
    double all = 0.0, rest = 0.0;
    if (E != INT32_MIN) {
        for (int64_t t = tid; t < tpr; t += 256) {
            const double key = tile_key(pm[t], pe[t]);
            if (key == -INFINITY) continue;
            const double v = ldexp(pm[t], pe[t] - E);
            all += v;
            if (before(lk, li, key, t)) rest += v;
        }
    }
    all = wave_sum_f64(all);
    rest = wave_sum_f64(rest);
    if (lane == 0) {
        s_s[wv] = all;
        s_k[wv] = rest;
    }
    __syncthreads();
    if (tid == 0) {
        const double A = s_s[0] + s_s[1] + s_s[2] + s_s[3];
        const double Rs = s_k[0] + s_k[1] + s_k[2] + s_k[3];
        frac[k] = A > 0.0 ? Rs / A : 0.0;
    }
}

// Each picked mixed tile's partial sum <- the sum of its FP64 tiles' partials.
__global__ void refine_merge_kernel(const int64_t* __restrict__ sel, int K, int ratio, int64_t tpr, int64_t tpr64,
                                    const double* __restrict__ p64_m, const int* __restrict__ p64_e,
                                    double* __restrict__ part_m, int* __restrict__ part_e)
{
    const int k = blockIdx.x;
    for (int r = threadIdx.x; r < K; r += blockDim.x) {
        const int64_t g = sel[(int64_t)k * K + r];
        if (g < 0) continue;
        const int64_t jt = g - (int64_t)k * tpr;
        int E = INT32_MIN;
        for (int q = 0; q < ratio; ++q) {
            const int64_t t64 = (int64_t)ratio * jt + q;
            if (t64 >= tpr64) break;
            const int64_t i64 = (int64_t)k * tpr64 + t64;
            if (p64_m[i64] != 0.0) E = max(E, p64_e[i64]);
        }
        double S = 0.0;
        if (E == INT32_MIN) {
            E = 0;
        } else {
            for (int q = 0; q < ratio; ++q) {
                const int64_t t64 = (int64_t)ratio * jt + q;
                if (t64 >= tpr64) break;
                const int64_t i64 = (int64_t)k * tpr64 + t64;
                if (p64_m[i64] != 0.0) S += ldexp(p64_m[i64], p64_e[i64] - E);
            }
        }
        part_m[g] = S;
        part_e[g] = E;
    }
}

hipError_t refine_select(const double* part_m, const int* part_e, int64_t tpr, int n_reps, int K, int ratio,
                         int64_t tpr64, int64_t* sel, int64_t* list64, double* frac, hipStream_t s)
{
    hipLaunchKernelGGL(refine_select_kernel, dim3(n_reps), dim3(256), 0, s, part_m, part_e, tpr, K, ratio, tpr64,
                       sel, list64, frac);
    return hipGetLastError();
}

hipError_t refine_merge(const int64_t* sel, int n_reps, int K, int ratio, int64_t tpr, int64_t tpr64,
                        const double* p64_m, const int* p64_e, double* part_m, int* part_e, hipStream_t s)
{
    hipLaunchKernelGGL(refine_merge_kernel, dim3(n_reps), dim3(64), 0, s, sel, K, ratio, tpr, tpr64, p64_m, p64_e,
                       part_m, part_e);
    return hipGetLastError();
}

}  // namespace dbg
