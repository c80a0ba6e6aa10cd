// Batched integration of many independent problems of uneven dimension in
// one persistent launch (no counterpart in the reference, which solves one
// problem per qmc_pnorm! call).
//
// Work items are (problem, shift, tile of 256 points), ordered largest d
// first; resident workgroups take the next item from a global counter until
// the queue is empty, so long and short problems interleave without a tail.
// Each item runs the dense integration (one sample per lane) over dimension
// blocks of W on its problem's factor: the block's conditional means from the
// earlier blocks, then walk_block_reg, which keeps them in registers.  For
// W = 8, 16 (and FP64) the contraction is register-blocked vector ALU work; for
// the mixed path at W = 32 it runs on the FP16 matrix cores in the split form
// of the dense split kernel (high and low FP16 parts of 2^6 y and 2^13 u,
// three products, FP32 accumulation), with the tiles transposed back to one
// sample per lane by lane-row swaps.  W is picked per batch by the autotuner
// of the C ABI layer; the grid is the kernel's resident workgroups
// (batch_occupancy).
#include <type_traits>

#include "dbg_device.h"

namespace dbg {

constexpr int kTileB = 256;
int batch_tile() { return kTileB; }

// queue head and status word back to zero (a kernel, not a memset node, so
// that the step can be captured in a graph)
__global__ void queue_reset_kernel(unsigned long long* head, int* status)
{
    *head = 0ull;
    *status = 0;
}

hipError_t queue_reset(const BatchArgs& B, hipStream_t s)
{
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(1), 0, s, B.counter, B.status);
    return hipGetLastError();
}

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef unsigned int uint2v __attribute__((ext_vector_type(2)));
constexpr int kYGB = 2 * kTileB * 8;          // halves per k-group of 8 dimensions in the split y slot
constexpr float kYScaleB = 64.0f;             // 2^6
constexpr float kUScaleB = 8192.0f;           // 2^13
constexpr float kAccScaleB = 1.0f / 524288.0f;  // 2^-19
constexpr float kHalfMaxB = 65504.0f;

// v -> (hi, lo), hi = fp16(v), lo = fp16(v - hi); |v| clamped to the FP16 range
__device__ __forceinline__ void split16b(float v, _Float16& hi, _Float16& lo)
{
    v = fabsf(v) > kHalfMaxB ? copysignf(kHalfMaxB, v) : v;
    hi = (_Float16)v;
    lo = (_Float16)(v - (float)hi);
}

// 4 x 4 transpose of (register, row of 16 lanes): afterwards register q at
// lane lr + 16 lq holds what register lq held at lane lr + 16 q
__device__ __forceinline__ void transpose_rows4(float& r0, float& r1, float& r2, float& r3)
{
    unsigned a0 = __float_as_uint(r0), a1 = __float_as_uint(r1), a2 = __float_as_uint(r2), a3 = __float_as_uint(r3);
    uint2v x = __builtin_amdgcn_permlane32_swap(a0, a2, false, false);
    a0 = x[0];
    a2 = x[1];
    x = __builtin_amdgcn_permlane32_swap(a1, a3, false, false);
    a1 = x[0];
    a3 = x[1];
    x = __builtin_amdgcn_permlane16_swap(a0, a1, false, false);
    a0 = x[0];
    a1 = x[1];
    x = __builtin_amdgcn_permlane16_swap(a2, a3, false, false);
    a2 = x[0];
    a3 = x[1];
    r0 = __uint_as_float(a0);
    r1 = __uint_as_float(a1);
    r2 = __uint_as_float(a2);
    r3 = __uint_as_float(a3);
}

}  // namespace

// The mixed item on the FP16 matrix cores (W a multiple of 32).  Layouts:
//   factor uc[j * ucs + 64 c + e] (j = target column < d + 32): e < 32 the high
//          parts of 2^13 U[32 c + e, j], e >= 32 the low parts; zero for rows
//          32 c + e >= j and for j >= d (batch_pack_split);
//   y      the workgroup's slot, k-group g of 8 dimensions: 256 samples x 8
//          high parts, then 256 x 8 low parts (the walk writes its sample's
//          block of W dimensions as W / 8 groups).
// The contraction of a block at g0 forms S^T (targets x samples) tiles: lane
// (lr, lq) supplies the A operand of target g0 + 1 + 16 tb + lr, rows 32 c +
// 8 lq .. + 7, and the B operand of sample 64 wv + 16 sb + lr, k-group 4 c +
// lq; each output lane holds targets 16 tb + 4 lq + r of sample 16 sb + lr,
// and transpose_rows4 over sb gives lane lr + 16 lq the targets of its own
// sample.  Each wave reads only its own 64 samples' y.
template <int W, int FORM>
__device__ __forceinline__ Weight<float> batch_item_split(const PointArgs& pts, const BatchProblem& P,
                                                          const double* f, int r, int64_t j, int tid,
                                                          _Float16* __restrict__ Ys16,
                                                          const _Float16* __restrict__ uc,
                                                          const float* __restrict__ Ucol,
                                                          const float* __restrict__ ak, const float* __restrict__ bk,
                                                          const float* __restrict__ invk)
{
    static_assert(W % 32 == 0, "whole chunks of 32 dimensions");
    using V = floatx4;
    constexpr int NTB = W / 16;
    const int lane = tid & 63, lr = lane & 15, lq = lane >> 4, wv = tid >> 6;
    const int d = P.d, nd = d - 1;
    const int64_t ucs = P.ucs;
    typename Path<true, FORM>::State st;
    Weight<float> w;
    Path<true, FORM>::init(st, w, f);
    for (int g0 = 0; g0 < nd; g0 += W) {
        V acc[NTB][4];
#pragma unroll
        for (int tb = 0; tb < NTB; ++tb)
#pragma unroll
            for (int sb = 0; sb < 4; ++sb) acc[tb][sb] = V{0, 0, 0, 0};
        const int nch = g0 / 32;
        const _Float16* __restrict__ ua = uc + (size_t)(g0 + 1 + lr) * ucs + 8 * lq;
        const _Float16* __restrict__ yb = Ys16 + (size_t)lq * kYGB + (size_t)(wv * 64 + lr) * 8;
        for (int c = 0; c < nch; ++c) {
            half8 ah[NTB], al[NTB], bh[4], bl[4];
#pragma unroll
            for (int tb = 0; tb < NTB; ++tb) {
                const _Float16* __restrict__ up = ua + (size_t)(16 * tb) * ucs + 64 * c;
                ah[tb] = *(const half8*)up;
                al[tb] = *(const half8*)(up + 32);
            }
#pragma unroll
            for (int sb = 0; sb < 4; ++sb) {
                const _Float16* __restrict__ yp = yb + (size_t)(4 * c) * kYGB + 16 * sb * 8;
                bh[sb] = *(const half8*)yp;
                bl[sb] = *(const half8*)(yp + kTileB * 8);
            }
#pragma unroll
            for (int tb = 0; tb < NTB; ++tb)
#pragma unroll
                for (int sb = 0; sb < 4; ++sb) {
                    acc[tb][sb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[tb], bh[sb], acc[tb][sb], 0, 0, 0);
                    acc[tb][sb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[tb], bl[sb], acc[tb][sb], 0, 0, 0);
                    acc[tb][sb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[tb], bh[sb], acc[tb][sb], 0, 0, 0);
                }
        }
        float s[W];
#pragma unroll
        for (int tb = 0; tb < NTB; ++tb)
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                float v0 = acc[tb][0][rr] * kAccScaleB, v1 = acc[tb][1][rr] * kAccScaleB;
                float v2 = acc[tb][2][rr] * kAccScaleB, v3 = acc[tb][3][rr] * kAccScaleB;
                transpose_rows4(v0, v1, v2, v3);
                s[16 * tb + rr] = v0;
                s[16 * tb + 4 + rr] = v1;
                s[16 * tb + 8 + rr] = v2;
                s[16 * tb + 12 + rr] = v3;
            }
        const int cnt = min(W, nd - g0);
        float yv[W];
        walk_block_reg<true, kGenRichtmyer, FORM, W, false>(pts, g0, cnt, r, j, tid, st, w, s, yv,
                                                            (float*)nullptr, 0, Ucol, d, ak, bk, invk);
        // the block's y, split, as W / 8 k-groups of this sample
#pragma unroll
        for (int h = 0; h < W / 8; ++h) {
            half8 hi, lo;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float v = (8 * h + e < cnt) ? yv[8 * h + e] * kYScaleB : 0.0f;
                _Float16 a, b;
                split16b(v, a, b);
                hi[e] = a;
                lo[e] = b;
            }
            _Float16* dst = Ys16 + (size_t)(g0 / 8 + h) * kYGB + (size_t)tid * 8;
            *(half8*)dst = hi;
            *(half8*)(dst + kTileB * 8) = lo;
        }
        // the next block's contraction reads these groups from other lanes of the wave
        __syncthreads();
    }
    return w;
}

// One work item's walk over its problem: the register-blocked contraction of
// each block of W dimensions into the lane's own registers, then the walk of
// the block from those registers (no LDS: occupancy is set by registers);
// FORM kPhiTailLower for a problem whose lower bounds are all -inf (mixed only)
template <bool MIXED, int W, int FORM, typename T>
__device__ __forceinline__ Weight<T> batch_item(const PointArgs& pts, const BatchProblem& P, const double* f,
                                                int r, int64_t j, int tid, T* __restrict__ Ys,
                                                const T* __restrict__ Uk, const T* __restrict__ Ucol,
                                                const T* __restrict__ ak, const T* __restrict__ bk,
                                                const T* __restrict__ invk)
{
    const int d = P.d, nd = d - 1;
    typename Path<MIXED, FORM>::State st;
    Weight<T> w;
    Path<MIXED, FORM>::init(st, w, f);
    for (int g0 = 0; g0 < nd; g0 += W) {
        T s[W];
#pragma unroll
        for (int c = 0; c < W; ++c) s[c] = T(0);
        for (int l = 0; l < g0; ++l) {
            const T y = Ys[(size_t)l * kTileB + tid];
            const T* __restrict__ ur = Uk + (size_t)l * P.ldu + g0 + 1;
#pragma unroll
            for (int c = 0; c < W; ++c) s[c] = fma(y, ur[c], s[c]);
        }
        T yv[W];
        walk_block_reg<MIXED, kGenRichtmyer, FORM, W>(pts, g0, min(W, nd - g0), r, j, tid, st, w, s, yv, Ys,
                                                      kTileB, Ucol, d, ak, bk, invk);
    }
    return w;
}

// the matrix-core form is held to four resident workgroups per CU (128
// registers): the walk, latency-bound, needs the waves more than the
// contraction needs registers for loads in flight
template <bool MIXED, int W, typename T = typename std::conditional<MIXED, float, double>::type>
__global__ __launch_bounds__(kTileB, (MIXED && W % 32 == 0) ? 4 : 1) void batch_integrate_kernel(BatchArgs B, const T* __restrict__ Ukall,
                                                                 const T* __restrict__ Ucall,
                                                                 const _Float16* __restrict__ ucall,
                                                                 T* __restrict__ Yall,
                                                                 double* __restrict__ part_m,
                                                                 int* __restrict__ part_e)
{
    __shared__ int64_t s_item;
    const int tid = threadIdx.x;
    T* __restrict__ Ys = Yall + (size_t)blockIdx.x * (size_t)B.y_slot;

    // Work-queue loop.  Thread 0 takes the next item in ONE region at the end
    // of the body (after storing the item's partial sum), closed by a barrier
    // before the whole workgroup reads the item and tests the loop condition.
    // A loop written with a thread-0 fetch at the top and a thread-0 store at
    // the bottom let the compiler join the two regions across the back-edge
    // into a nested loop that only thread 0 left after an item: the other
    // threads re-read the old item forever.  Each fetched item is strictly
    // larger than the previous one; a repeat stops the workgroup and is
    // reported through the status word instead of spinning.
    if (tid == 0) s_item = (int64_t)atomicAdd(B.counter, 1ull);
    __syncthreads();
    int64_t it = s_item;
    int64_t prev = -1;
    while (it < B.n_items) {  // uniform: it comes from one LDS word after a barrier
        if (it <= prev) {
            if (tid == 0) atomicOr(B.status, 1);
            break;
        }
        prev = it;
        __syncthreads();  // every thread has read s_item before thread 0 rewrites it
        const int2 wi = B.items[it];
        const int p = wi.x;
        const int lt = wi.y;
        const BatchProblem P = B.probs[p];
        const int nd = P.d - 1;
        const int r = lt / P.tiles_per_rep;
        const int t = lt - r * P.tiles_per_rep;
        const int64_t j = (int64_t)t * kTileB + tid + 1;
        const bool valid = j <= B.m;
        const T* __restrict__ Uk = Ukall + P.uk_off;
        const T* __restrict__ Ucol = Ucall + P.ucol_off;
        const T* __restrict__ ak = (const T*)B.ak + P.vec_off;
        const T* __restrict__ bk = (const T*)B.bk + P.vec_off;
        const T* __restrict__ invk = (const T*)B.invk + P.vec_off;
        const double* f = B.first + 8 * (int64_t)p;
        PointArgs pts{};
        pts.nd = nd;
        pts.q = B.q;
        pts.shifts_f = B.shifts + P.shift_off;
        pts.n_base = 0;
        pts.q64 = B.q64;
        pts.shift32 = B.shift32 + P.shift_off;

        Weight<T> w;
        if constexpr (MIXED && W % 32 == 0) {
            _Float16* Ys16 = reinterpret_cast<_Float16*>(Ys);
            const _Float16* uc = ucall + P.uc_off;
            if (P.lower)  // uniform: one problem per item
                w = batch_item_split<W, kPhiTailLower>(pts, P, f, r, j, tid, Ys16, uc, Ucol, ak, bk, invk);
            else
                w = batch_item_split<W, kPhiErf>(pts, P, f, r, j, tid, Ys16, uc, Ucol, ak, bk, invk);
        } else if (MIXED && P.lower) {
            w = batch_item<MIXED, W, MIXED ? kPhiTailLower : kPhiErf>(pts, P, f, r, j, tid, Ys, Uk, Ucol, ak, bk,
                                                                      invk);
        } else {
            w = batch_item<MIXED, W, kPhiErf>(pts, P, f, r, j, tid, Ys, Uk, Ucol, ak, bk, invk);
        }
        const double pm = valid ? (double)w.pm : 0.0;
        double S;
        int E;
        block_scaled_sum<kTileB>(pm, w.pe, S, E);
        if (tid == 0) {
            part_m[P.part_off + lt] = S;
            part_e[P.part_off + lt] = E;
            s_item = (int64_t)atomicAdd(B.counter, 1ull);
        }
        __syncthreads();
        it = s_item;
    }
}

template <bool MIXED, int W>
static void launch_bi(const BatchArgs& B, int grid, const void* Uk, const void* Ucol, const void* uc16, void* Y,
                      double* pm, int* pe, hipStream_t s)
{
    using T = typename std::conditional<MIXED, float, double>::type;
    hipLaunchKernelGGL((batch_integrate_kernel<MIXED, W>), dim3(grid), dim3(kTileB), 0, s, B,
                       (const T*)Uk, (const T*)Ucol, (const _Float16*)uc16, (T*)Y, pm, pe);
}

template <bool MIXED, int W>
static int occ_bi()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_integrate_kernel<MIXED, W>, kTileB, 0) != hipSuccess)
        return 0;
    return n;
}

// resident workgroups per CU of the integration kernel at width W (0 on error)
int batch_occupancy(bool mixed, int W)
{
    if (mixed) return W == 8 ? occ_bi<true, 8>() : (W == 32 ? occ_bi<true, 32>() : occ_bi<true, 16>());
    return W == 8 ? occ_bi<false, 8>() : (W == 32 ? occ_bi<false, 32>() : occ_bi<false, 16>());
}

hipError_t batch_integrate(const BatchArgs& B, bool mixed, int W, int grid, const void* Uk,
                           const void* Ucol, const void* uc16, void* Y, double* part_m, int* part_e, hipStream_t s)
{
    if (mixed) {
        if (W == 8) launch_bi<true, 8>(B, grid, Uk, Ucol, uc16, Y, part_m, part_e, s);
        else if (W == 32) launch_bi<true, 32>(B, grid, Uk, Ucol, uc16, Y, part_m, part_e, s);
        else launch_bi<true, 16>(B, grid, Uk, Ucol, uc16, Y, part_m, part_e, s);
    } else {
        if (W == 8) launch_bi<false, 8>(B, grid, Uk, Ucol, uc16, Y, part_m, part_e, s);
        else if (W == 32) launch_bi<false, 32>(B, grid, Uk, Ucol, uc16, Y, part_m, part_e, s);
        else launch_bi<false, 16>(B, grid, Uk, Ucol, uc16, Y, part_m, part_e, s);
    }
    return hipGetLastError();
}

// One block per problem: per-shift means of the tile partials (Neumaier sums
// in FP64 with a common exponent), then mean and 3*sd over the shifts.
__global__ __launch_bounds__(256) void batch_reduce_kernel(BatchArgs B, const double* __restrict__ part_m,
                                                           const int* __restrict__ part_e)
{
    __shared__ int s_e[4];
    __shared__ double s_s[256], s_c[256];
    const int p = blockIdx.x;
    const int tid = threadIdx.x;
    const BatchProblem P = B.probs[p];
    const int R = B.n_reps;
    double* rm = B.rep_m + (size_t)p * R;
    int* re = B.rep_e + (size_t)p * R;
    for (int r = 0; r < R; ++r) {
        const double* pm = part_m + P.part_off + (int64_t)r * P.tiles_per_rep;
        const int* pe = part_e + P.part_off + (int64_t)r * P.tiles_per_rep;
        int e = INT32_MIN;
        for (int t = tid; t < P.tiles_per_rep; t += 256)
            if (pm[t] != 0.0) e = max(e, pe[t]);
        e = wave_max_int(e);
        if ((tid & 63) == 0) s_e[tid >> 6] = e;
        __syncthreads();
        int E = max(max(s_e[0], s_e[1]), max(s_e[2], s_e[3]));
        if (E == INT32_MIN) E = 0;
        double sum = 0.0, comp = 0.0;
        for (int t = tid; t < P.tiles_per_rep; t += 256) {
            const double v = pm[t] != 0.0 ? ldexp(pm[t], pe[t] - E) : pm[t];
            const double tt = sum + v;
            comp += (fabs(sum) >= fabs(v)) ? (sum - tt) + v : (v - tt) + sum;
            sum = tt;
        }
        s_s[tid] = sum;
        s_c[tid] = comp;
        __syncthreads();
        if (tid == 0) {
            double S = 0.0, Cc = 0.0;
            for (int q = 0; q < 256; ++q) {
                const double v = s_s[q];
                const double tt = S + v;
                Cc += (fabs(S) >= fabs(v)) ? (S - tt) + v : (v - tt) + S;
                S = tt;
                Cc += s_c[q];
            }
            S = (S + Cc) * (1.0 / (double)B.m);
            int en = 0;
            const double mn = frexp(S, &en);
            rm[r] = mn;
            re[r] = (mn == 0.0 || !isfinite(mn)) ? 0 : E + en;
        }
        __syncthreads();
    }
    if (tid == 0) {
        int E = INT32_MIN;
        for (int r = 0; r < R; ++r)
            if (rm[r] != 0.0 && re[r] > E) E = re[r];
        if (E == INT32_MIN) E = 0;
        double sum = 0.0;
        for (int r = 0; r < R; ++r) sum += ldexp(rm[r], re[r] - E);
        const double result = sum / R;
        double acc = 0.0;
        for (int r = 0; r < R; ++r) {
            const double v = ldexp(rm[r], re[r] - E) - result;
            acc += v * v / R;
        }
        const double err = 3.0 * sqrt(acc / (R - 1));
        double* o = B.out + 8 * (int64_t)p;
        const bool ok = B.first[8 * (int64_t)p + 6] != 0.0;
        o[0] = ok ? ldexp(result, E) : NAN;
        o[1] = ok ? (result > 0 ? log(result) + E * 0.69314718055994530942 : (result == 0 ? -INFINITY : NAN)) : NAN;
        o[2] = ldexp(err, E);
        o[3] = err / result;
        o[4] = result;
        o[5] = err;
        o[6] = (double)E;
        o[7] = (double)B.rank[p];
    }
}

hipError_t batch_reduce(const BatchArgs& B, const double* part_m, const int* part_e, hipStream_t s)
{
    hipLaunchKernelGGL(batch_reduce_kernel, dim3(B.K), dim3(256), 0, s, B, part_m, part_e);
    return hipGetLastError();
}

// The split-FP16 factor of every problem (mixed batch): one workgroup per
// problem writes uc[j * ucs + 64 c + e] for all j < d + 32 and rows 32 c + e <
// ucs / 2 from the FP32 column-major factor (ldc = d): the high and low FP16
// parts of 2^13 U[32 c + e, j] for 32 c + e < j < d, else zero.
__global__ __launch_bounds__(256) void batch_pack_split_kernel(BatchArgs B, const float* __restrict__ Ucall,
                                                               _Float16* __restrict__ ucall)
{
    const BatchProblem P = B.probs[blockIdx.x];
    const int d = P.d;
    const int64_t ucs = P.ucs, rows = ucs / 2;
    const float* __restrict__ Ucol = Ucall + P.ucol_off;
    _Float16* __restrict__ uc = ucall + P.uc_off;
    const int64_t total = (int64_t)(d + 32) * rows;
    for (int64_t idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int jc = (int)(idx / rows);
        const int l = (int)(idx - (int64_t)jc * rows);
        const float v = (jc < d && l < jc) ? Ucol[(size_t)jc * d + l] * kUScaleB : 0.0f;
        _Float16 hi, lo;
        split16b(v, hi, lo);
        const int c = l / 32, e = l % 32;
        uc[(size_t)jc * ucs + 64 * c + e] = hi;
        uc[(size_t)jc * ucs + 64 * c + 32 + e] = lo;
    }
}

hipError_t batch_pack_split(const BatchArgs& B, const void* Ucol, void* uc16, hipStream_t s)
{
    hipLaunchKernelGGL(batch_pack_split_kernel, dim3(B.K), dim3(256), 0, s, B, (const float*)Ucol, (_Float16*)uc16);
    return hipGetLastError();
}

}  // namespace dbg
