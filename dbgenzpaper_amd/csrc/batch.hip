// Batched integration of many independent problems of uneven dimension in
// one persistent launch (no counterpart in the reference, which solves one
// problem per qmc_pnorm! call).
//
// Work items are (problem, shift, tile of 256 points), ordered largest d
// first; resident workgroups take the next item from a global counter until
// the queue is empty, so long and short problems interleave without a tail.
// Each item runs the vector-ALU dense integration (one sample per lane,
// dimension blocks of W with the register-blocked contraction, then
// walk_block_reg, which keeps the hand-off in registers) on its problem's
// factor.  W is a template parameter picked per batch by the autotuner of the
// C ABI layer; the grid is the kernel's resident workgroups (batch_occupancy).
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
        walk_block_reg<MIXED, kGenRichtmyer, FORM, W>(pts, g0, min(W, nd - g0), r, j, tid, st, w, s, Ys, kTileB,
                                                      Ucol, d, ak, bk, invk);
    }
    return w;
}

template <bool MIXED, int W, typename T = typename std::conditional<MIXED, float, double>::type>
__global__ __launch_bounds__(kTileB) void batch_integrate_kernel(BatchArgs B, const T* __restrict__ Ukall,
                                                                 const T* __restrict__ Ucall,
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
        if (MIXED && P.lower)  // uniform: one problem per item
            w = batch_item<MIXED, W, MIXED ? kPhiTailLower : kPhiErf>(pts, P, f, r, j, tid, Ys, Uk, Ucol, ak, bk,
                                                                      invk);
        else
            w = batch_item<MIXED, W, kPhiErf>(pts, P, f, r, j, tid, Ys, Uk, Ucol, ak, bk, invk);
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
static void launch_bi(const BatchArgs& B, int grid, const void* Uk, const void* Ucol, void* Y,
                      double* pm, int* pe, hipStream_t s)
{
    using T = typename std::conditional<MIXED, float, double>::type;
    hipLaunchKernelGGL((batch_integrate_kernel<MIXED, W>), dim3(grid), dim3(kTileB), 0, s, B,
                       (const T*)Uk, (const T*)Ucol, (T*)Y, pm, pe);
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
                           const void* Ucol, void* Y, double* part_m, int* part_e, hipStream_t s)
{
    if (mixed) {
        if (W == 8) launch_bi<true, 8>(B, grid, Uk, Ucol, Y, part_m, part_e, s);
        else if (W == 32) launch_bi<true, 32>(B, grid, Uk, Ucol, Y, part_m, part_e, s);
        else launch_bi<true, 16>(B, grid, Uk, Ucol, Y, part_m, part_e, s);
    } else {
        if (W == 8) launch_bi<false, 8>(B, grid, Uk, Ucol, Y, part_m, part_e, s);
        else if (W == 32) launch_bi<false, 32>(B, grid, Uk, Ucol, Y, part_m, part_e, s);
        else launch_bi<false, 16>(B, grid, Uk, Ucol, Y, part_m, part_e, s);
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

}  // namespace dbg
