// FP64 Cholesky factorizations on the device.
//
// cholesky_genz!  (src/cholesky.jl:445-628) and cholesky_classic!
// (:144-242).  The reference's four-level blocking only changes rounding,
// with one exception that changes results: the Genz-Bretz pivot is searched
// inside the current bs2-window ("ii" block) only, and the conditional means
// s_vec only collect contributions of pivots of the same window.  The device
// version therefore factors window by window:
//
//   chol_window_kernel  one workgroup factors the window's diagonal block with
//                       the block-local pivot (argmin of erf(b')-erf(a'),
//                       first index, NaN first), swapping rows and columns of
//                       the whole upper-stored matrix, and the window's rows
//                       of U inside the window; inner blocks of nb pivots are
//                       left-looking with a deferred trailing update, like
//                       the reference's bs1 blocks (d_vec reset per block);
//   TRSM (rocBLAS)      the window's rows of U beyond the window
//                       (ldiv!(UpperTriangular(.)', .) of :589-623);
//   chol_syrk_kernel    the trailing Schur complement updates (the mul! of
//                       :578-582 inside a window and :592-595 after it), upper
//                       triangle only.
//
// The matrix is column-major with leading dimension lda.  Only the upper
// triangle is kept up to date; chol_finish zeroes the strict lower triangle.
#include <float.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <type_traits>

#include "dbg_device.h"

namespace dbg {

constexpr int kCholThreads = 1024;
constexpr int kMaxWin = 2048;

int chol_max_window() { return kMaxWin; }

#define MM(i, j) M[(size_t)(i) + (size_t)(j) * (size_t)lda]

// scale_j = sqrt(M_jj) (1 if M_jj <= 0), eps0 = eps * max|diag| of the input
// (src/cholesky.jl:450-456), rank = n.
__global__ __launch_bounds__(kCholThreads) void chol_diag_kernel(const double* M, int64_t lda, int n,
                                                                 double* scale, double* eps0,
                                                                 int32_t* rank)
{
    __shared__ double s_m[kCholThreads / 64];
    double mx = 0.0;
    for (int j = threadIdx.x; j < n; j += kCholThreads) {
        const double v = MM(j, j);
        const double av = fabs(v);
        mx = (isnan(av) || av > mx) ? av : mx;
        scale[j] = v > 0 ? sqrt(v) : 1.0;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(mx, off, 64);
        mx = (isnan(o) || o > mx) ? o : mx;
    }
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = s_m[0];
        for (int w = 1; w < kCholThreads / 64; ++w) m = (isnan(s_m[w]) || s_m[w] > m) ? s_m[w] : m;
        *eps0 = DBL_EPSILON * m;
        *rank = n;
    }
}

// Standardize to a correlation matrix and scale the bounds
// (src/cholesky.jl:458-473).
__global__ void chol_scale_kernel(double* M, int64_t lda, int n, const double* __restrict__ scale,
                                  double* a, double* b, int mirror)
{
    const int j = blockIdx.y;
    const double inv_s_j = 1.0 / scale[j];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= j; i += gridDim.x * blockDim.x) {
        if (i == j) {
            MM(j, j) = 1.0;
            a[j] *= inv_s_j;
            b[j] *= inv_s_j;
        } else {
            const double val = MM(i, j) * (inv_s_j / scale[i]);
            MM(i, j) = val;
            if (mirror) MM(j, i) = val;
        }
    }
}

hipError_t chol_standardize(const CholArgs& C, bool mirror, hipStream_t s)
{
    hipLaunchKernelGGL(chol_diag_kernel, dim3(1), dim3(kCholThreads), 0, s, C.M, C.lda, C.n,
                       C.scale, C.eps0, C.rank);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    dim3 g((C.n + 255) / 256, C.n);
    hipLaunchKernelGGL(chol_scale_kernel, g, dim3(256), 0, s, C.M, C.lda, C.n, C.scale, C.a, C.b,
                       mirror ? 1 : 0);
    return hipGetLastError();
}

// acc - sum_{l = l0}^{l1 - 1} M[l, x] M[l, y], subtracted in ascending l (the
// order of the reference's loops); the loads go out in batches of eight so a
// batch costs one memory round trip
__device__ __forceinline__ double sub_dot(const double* M, int64_t lda, double acc, int l0, int l1, int x, int y)
{
    int l = l0;
    for (; l + 8 <= l1; l += 8) {
        double a[8], b[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            a[q] = MM(l + q, x);
            b[q] = MM(l + q, y);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) acc -= a[q] * b[q];
    }
    for (; l < l1; ++l) acc -= MM(l, x) * MM(l, y);
    return acc;
}

// argmin key: NaN before any number, then smaller value, then smaller index
__device__ __forceinline__ bool better(double v, int i, double bv, int bi)
{
    const bool vn = isnan(v), bn = isnan(bv);
    if (vn != bn) return vn;
    if (!vn && v != bv) return v < bv;
    return i < bi;
}

// Pivot modes of the window kernels
constexpr int kPivotNone = 0;    // cholesky_classic!
constexpr int kPivotGenz = 1;    // cholesky_genz!: smallest interval mass
constexpr int kPivotRowMax = 2;  // cholesky_rowmax!: largest Schur diagonal

// argmax key of cholesky_rowmax! (src/cholesky.jl:319-327: a scan from i with
// a strict >, so the first maximum wins and NaN is never taken after i): a
// NaN loses to every number, then larger value, then smaller index.  A NaN at
// i itself keeps i; the caller checks that case.
__device__ __forceinline__ bool better_max(double v, int i, double bv, int bi)
{
    const bool vn = isnan(v), bn = isnan(bv);
    if (vn != bn) return bn;
    if (!vn && v != bv) return v > bv;
    return i < bi;
}

// pivot selection shared by both window kernels: reduce (bv, bi) over a
// wavefront by the mode's key
__device__ __forceinline__ void pivot_wave_reduce(int pivot, double& bv, int& bi)
{
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(bv, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        const bool take = pivot == kPivotRowMax ? better_max(ov, oi, bv, bi) : better(ov, oi, bv, bi);
        if (take) {
            bv = ov;
            bi = oi;
        }
    }
}

__device__ __forceinline__ bool pivot_better(int pivot, double v, int i, double bv, int bi)
{
    return pivot == kPivotRowMax ? better_max(v, i, bv, bi) : better(v, i, bv, bi);
}

// Factor the pivots [ib_begin, ib_end) of the window [w0, w1) (whole inner
// blocks of nb) with one workgroup of NT threads (window size at most WMAX).
// With inkernel_update the deferred update of the rest of the window after
// each inner block is done here; otherwise the caller launches it.  On a rank
// failure *C.rank is set and the function returns.
template <int NT, int WMAX>
__device__ __forceinline__ void chol_window_body(const CholArgs& C, int w0, int w1, int ib_begin,
                                                 int ib_end, int nb, int pivot, double eps0,
                                                 bool inkernel_update)
{
    constexpr int kCholThreads = NT;
    constexpr int kMaxWin = WMAX;
    __shared__ double s_dsum[kMaxWin], s_sv[kMaxWin], s_a[kMaxWin], s_b[kMaxWin];
    __shared__ int s_perm[kMaxWin];
    __shared__ double s_rv[kCholThreads / 64];
    __shared__ int s_ri[kCholThreads / 64];
    __shared__ double s_uii, s_yi;
    __shared__ int s_im, s_stop;

    double* M = C.M;
    const int64_t lda = C.lda;
    const int n = C.n;
    const int tid = threadIdx.x;
    const double c2 = 1.0 / sqrt(2.0);
    const double is2pi = 2.0 / sqrt(2.0 * 3.14159265358979323846);

    for (int k = tid; k < w1 - w0; k += kCholThreads) {
        s_sv[k] = C.s_vec[w0 + k];
        s_a[k] = C.a[w0 + k];
        s_b[k] = C.b[w0 + k];
        s_perm[k] = C.perm[w0 + k];
    }
    if (tid == 0) s_stop = 0;
    __syncthreads();

    for (int ib0 = ib_begin; ib0 < ib_end; ib0 += nb) {
        const int ib1 = min(w1, ib0 + nb);
        for (int k = ib0 + tid; k < w1; k += kCholThreads) s_dsum[k - w0] = 0.0;
        __syncthreads();
        for (int i = ib0; i < ib1; ++i) {
            // ---- pivot search over the rest of the window (src/cholesky.jl:500-516)
            double bv = NAN;
            int bi = INT32_MAX;
            double best_order = 0.0;
            if (pivot) {
                bv = pivot == kPivotRowMax ? -INFINITY : INFINITY;
                for (int k = i + tid; k < w1; k += kCholThreads) {
                    double o;
                    if (pivot == kPivotRowMax) {
                        o = MM(k, k) - s_dsum[k - w0];
                    } else {
                        double m_kk = MM(k, k) - s_dsum[k - w0];
                        m_kk = m_kk < 0 ? 0.0 : sqrt(m_kk);
                        const double a_k = (s_a[k - w0] - s_sv[k - w0]) / m_kk * c2;
                        const double b_k = (s_b[k - w0] - s_sv[k - w0]) / m_kk * c2;
                        o = erf(b_k) - erf(a_k);
                    }
                    if (pivot_better(pivot, o, k, bv, bi)) {
                        bv = o;
                        bi = k;
                    }
                }
                pivot_wave_reduce(pivot, bv, bi);
                if ((tid & 63) == 0) {
                    s_rv[tid >> 6] = bv;
                    s_ri[tid >> 6] = bi;
                }
                __syncthreads();
            }
            if (tid == 0) {
                int im = i;
                if (pivot) {
                    bv = s_rv[0];
                    bi = s_ri[0];
                    for (int w = 1; w < kCholThreads / 64; ++w)
                        if (pivot_better(pivot, s_rv[w], s_ri[w], bv, bi)) {
                            bv = s_rv[w];
                            bi = s_ri[w];
                        }
                    im = bi;
                    best_order = bv;
                    // rowmax: a NaN at i keeps i (nothing compares greater)
                    if (pivot == kPivotRowMax && (bi == INT32_MAX || isnan(MM(i, i) - s_dsum[i - w0]))) im = i;
                }
                double m_ii = MM(im, im) - s_dsum[im - w0];
                // rank: genz stops at m_ii <= eps0 after the swap (src/cholesky.jl:518-541),
                // rowmax at m_ii <= eps0 before it, eps0 then being the rank tolerance
                // (0 for the reference's m_ii <= 0, :329-350)
                const bool stop = !(m_ii > eps0) && (pivot == kPivotGenz || pivot == kPivotRowMax);
                if (stop && pivot == kPivotRowMax) im = i;
                if (im != i) {
                    const int a0 = i - w0, a1 = im - w0;
                    int tp = s_perm[a0]; s_perm[a0] = s_perm[a1]; s_perm[a1] = tp;
                    double t;
                    t = s_dsum[a0]; s_dsum[a0] = s_dsum[a1]; s_dsum[a1] = t;
                    t = s_sv[a0]; s_sv[a0] = s_sv[a1]; s_sv[a1] = t;
                    if (pivot != kPivotRowMax) {  // rowmax leaves a and b in place
                        t = s_a[a0]; s_a[a0] = s_a[a1]; s_a[a1] = t;
                        t = s_b[a0]; s_b[a0] = s_b[a1]; s_b[a1] = t;
                    }
                }
                s_im = im;
                if (stop) {
                    s_stop = 1;
                    *C.rank = i;
                } else {
                    const double uii = sqrt(m_ii);
                    s_uii = uii;
                    const double a_norm = (s_a[i - w0] - s_sv[i - w0]) / uii;
                    const double b_norm = (s_b[i - w0] - s_sv[i - w0]) / uii;
                    s_yi = pivot == kPivotGenz
                               ? (exp(-(a_norm * a_norm) / 2) - exp(-(b_norm * b_norm) / 2)) * is2pi / best_order
                               : 0.0;
                }
            }
            __syncthreads();
            const int im = s_im;
            // ---- symmetric row/column swap i <-> im of the upper-stored matrix
            if (im != i) {
                // rows r < i: columns i and im
                for (int r = tid; r < i; r += kCholThreads) {
                    const double t = MM(r, i);
                    MM(r, i) = MM(r, im);
                    MM(r, im) = t;
                }
                // i < r < im: M[i, r] <-> M[r, im]
                for (int r = i + 1 + tid; r < im; r += kCholThreads) {
                    const double t = MM(i, r);
                    MM(i, r) = MM(r, im);
                    MM(r, im) = t;
                }
                // c > im: M[i, c] <-> M[im, c]
                for (int c = im + 1 + tid; c < n; c += kCholThreads) {
                    const double t = MM(i, c);
                    MM(i, c) = MM(im, c);
                    MM(im, c) = t;
                }
                if (tid == 0) {
                    const double t = MM(i, i);
                    MM(i, i) = MM(im, im);
                    MM(im, im) = t;
                }
            }
            __syncthreads();
            if (s_stop) break;
            const double uii = s_uii;
            const double yi = s_yi;
            if (tid == 0) MM(i, i) = uii;
            // ---- row i of U inside the window (src/cholesky.jl:547-570)
            for (int c = i + 1 + tid; c < w1; c += kCholThreads) {
                const double acc = sub_dot(M, lda, MM(i, c), ib0, i, i, c);
                const double u = acc / uii;
                MM(i, c) = u;
                s_dsum[c - w0] += u * u;
                s_sv[c - w0] += yi * u;
            }
            __syncthreads();
        }
        if (s_stop) break;
        // ---- deferred update of the rest of the window (src/cholesky.jl:578-582)
        if (inkernel_update && ib1 < w1) {
            const int wr = w1 - ib1;
            for (int64_t idx = tid; idx < (int64_t)wr * wr; idx += kCholThreads) {
                const int x = ib1 + (int)(idx % wr);
                const int y = ib1 + (int)(idx / wr);
                if (x <= y) MM(x, y) = sub_dot(M, lda, MM(x, y), ib0, ib1, x, y);
            }
            __syncthreads();
        }
    }
    for (int k = tid; k < w1 - w0; k += kCholThreads) {
        C.s_vec[w0 + k] = s_sv[k];
        C.a[w0 + k] = s_a[k];
        C.b[w0 + k] = s_b[k];
        C.perm[w0 + k] = s_perm[k];
    }
    __syncthreads();
}

// One inner block of nb pivots of the window [w0, w1).
__global__ __launch_bounds__(kCholThreads) void chol_window_kernel(CholArgs C, int w0, int w1, int ib0,
                                                                   int nb, int pivot)
{
    if (*C.rank < C.n) return;  // an earlier pivot found the matrix rank deficient
    chol_window_body<kCholThreads, kMaxWin>(C, w0, w1, ib0, min(w1, ib0 + nb), nb, pivot, *C.eps0, false);
}

// ---------------------------------------------------------------------------
// The same inner block with the window's columns in registers: thread t owns
// column c = ib0 + t of [ib0, w1) (at most kRegCols) and keeps the rows of U
// this block has produced for it, U[ib0 + q, c] for q < i - ib0, in ur[q].
// A pivot step then needs global memory only for the symmetric swap (which
// also hands each thread the new M[i, c]) and the stores of row i: the
// window's diagonal sits in LDS with the other per-index state, and the
// in-block dot of row i (src/cholesky.jl:547-570) multiplies the broadcast
// column i (LDS) with the thread's own registers, in the same ascending
// order and the same operations as chol_window_body, so the results and the
// pivot order are identical.
// ---------------------------------------------------------------------------
constexpr int kRegCols = 512;  // threads = window columns from the block start
constexpr int kRegNb = 64;     // rows per inner block held in registers

// Register layout: the block's rows for the thread's column in groups of 16
// pivots.  The group being filled is `cur`; completed groups move to g0..g2.
// The pivot loop is unrolled over the 16 positions of a group, so every
// register index is a compile-time constant (a runtime index would put the
// array in scratch memory).
struct ColRegs {
    double g0[16], g1[16], g2[16], cur[16];
};

// dst[0 .. 16 ng + E) = the column's rows (ng completed groups, then E of cur)
template <int E>
__device__ __forceinline__ void regs_store(const ColRegs& r, int ng, double* dst)
{
    if (ng > 0) {
#pragma unroll
        for (int e = 0; e < 16; ++e) dst[e] = r.g0[e];
    }
    if (ng > 1) {
#pragma unroll
        for (int e = 0; e < 16; ++e) dst[16 + e] = r.g1[e];
    }
    if (ng > 2) {
#pragma unroll
        for (int e = 0; e < 16; ++e) dst[32 + e] = r.g2[e];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) dst[16 * ng + e] = r.cur[e];
}

template <int E>
__device__ __forceinline__ void regs_load(ColRegs& r, int ng, const double* src)
{
    if (ng > 0) {
#pragma unroll
        for (int e = 0; e < 16; ++e) r.g0[e] = src[e];
    }
    if (ng > 1) {
#pragma unroll
        for (int e = 0; e < 16; ++e) r.g1[e] = src[16 + e];
    }
    if (ng > 2) {
#pragma unroll
        for (int e = 0; e < 16; ++e) r.g2[e] = src[32 + e];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) r.cur[e] = src[16 * ng + e];
}

// acc - sum_{l < 16 ng + E} ci[l] U[l], subtracted in ascending l
template <int E>
__device__ __forceinline__ double regs_dot(const ColRegs& r, int ng, const double* ci, double acc)
{
    if (ng > 0) {
#pragma unroll
        for (int e = 0; e < 16; ++e) acc -= ci[e] * r.g0[e];
    }
    if (ng > 1) {
#pragma unroll
        for (int e = 0; e < 16; ++e) acc -= ci[16 + e] * r.g1[e];
    }
    if (ng > 2) {
#pragma unroll
        for (int e = 0; e < 16; ++e) acc -= ci[32 + e] * r.g2[e];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) acc -= ci[16 * ng + e] * r.cur[e];
    return acc;
}

struct RegShared {
    double dsum[kRegCols], sv[kRegCols], a[kRegCols], b[kRegCols], dg[kRegCols];
    int perm[kRegCols];
    double ci[kRegNb], xch[kRegNb];
    double rv[kRegCols / 64];
    int ri[kRegCols / 64];
    // a pivot's swap of two indices' entries: [0] index i's, [1] index im's
    double x_dsum[2], x_sv[2], x_a[2], x_b[2], x_dg[2];
    int x_perm[2];
};

// a 64-bit lane value moved by a DPP control (rows of 16 lanes and the row
// broadcasts); lanes without a source, or outside row_mask, keep their own
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v)
{
    const int lo = (int)(uint32_t)v, hi = (int)(uint32_t)(v >> 32);
    const int l = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xf, false);
    const int h = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xf, false);
    return ((uint64_t)(uint32_t)h << 32) | (uint32_t)l;
}

// minimum of a 64-bit key over the (full) wavefront, in every lane: inclusive
// row minima (row_shr 1, 2, 4, 8), rows 1 and 3 take row 0 and 2 (row_bcast
// 15), rows 2 and 3 take row 1 (row_bcast 31); lane 63 then holds the minimum
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t k)
{
    k = min(k, dpp_u64<0x111, 0xf>(k));
    k = min(k, dpp_u64<0x112, 0xf>(k));
    k = min(k, dpp_u64<0x114, 0xf>(k));
    k = min(k, dpp_u64<0x118, 0xf>(k));
    k = min(k, dpp_u64<0x142, 0xa>(k));
    k = min(k, dpp_u64<0x143, 0xc>(k));
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)k, 63);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(k >> 32), 63);
    return ((uint64_t)hi << 32) | lo;
}

// An unsigned key whose order is the pivot mode's preference on one
// candidate's value: for the argmin of cholesky_genz! NaN first, then
// ascending; for the argmax of cholesky_rowmax! descending (NaN is never a
// candidate there).  -0 and +0 get one key, as they compare equal; equal keys
// are then decided by the smaller index.
__device__ __forceinline__ uint64_t pivot_key(int pivot, double v)
{
    if (isnan(v)) return 0;
    const uint64_t b = (uint64_t)__double_as_longlong(v + 0.0);  // -0 -> +0
    const uint64_t asc = (b >> 63) ? ~b : (b | 0x8000000000000000ull);
    return pivot == kPivotRowMax ? ~asc : asc;
}

// One pivot i = ib0 + 16 ng + E of the inner block; returns false when the
// block ends (rank deficiency).  All branches are uniform.  Two barriers per
// pivot: after the search (each wavefront's best candidate in LDS) and after
// the swap.  Every thread then reduces the wavefront minima and evaluates the
// pivot's diagonal and conditional mean itself (the same operations on the
// same LDS values, so all threads agree), and the per-index state in LDS is
// only ever written by the thread that owns the index: the two threads of a
// swap read each other's entries before the second barrier and write their
// own after it, when nobody else reads them until the next pivot's first.
template <int PIVOT, int E>
__device__ __forceinline__ bool reg_pivot(const CholArgs& C, RegShared& S, ColRegs& R, int ng, int ib0, int w1,
                                          double eps0)
{
    constexpr int pivot = PIVOT;
    double* M = C.M;
    const int64_t lda = C.lda;
    const int n = C.n;
    const int tid = threadIdx.x;
    const int q = 16 * ng + E;
    const int i = ib0 + q;
    const int c = ib0 + tid;
    const bool own = c < w1;
    const int a0 = i - ib0;
    const double c2 = 1.0 / sqrt(2.0);
    const double is2pi = 2.0 / sqrt(2.0 * 3.14159265358979323846);
    const double none_v = pivot == kPivotRowMax ? -INFINITY : INFINITY;
    // ---- pivot search over the rest of the window (src/cholesky.jl:500-516)
    if (pivot) {
        double o = 0.0;
        bool cand = false;
        if (own && c >= i) {
            if (pivot == kPivotRowMax) {
                o = S.dg[tid] - S.dsum[tid];
            } else {
                double m_kk = S.dg[tid] - S.dsum[tid];
                m_kk = m_kk < 0 ? 0.0 : sqrt(m_kk);
                const double a_k = (S.a[tid] - S.sv[tid]) / m_kk * c2;
                const double b_k = (S.b[tid] - S.sv[tid]) / m_kk * c2;
                o = erf(b_k) - erf(a_k);
            }
            // rowmax takes no NaN (src/cholesky.jl:319-327: a strict > from i)
            cand = pivot != kPivotRowMax || !isnan(o);
        }
        // the wavefront's best: the smallest key, then the smallest index
        // (lane) among the candidates that hold it
        const uint64_t key = cand ? pivot_key(pivot, o) : ~0ull;
        const uint64_t kmin = wave_min_u64(key);
        const uint64_t hit = __ballot(cand && key == kmin);
        if ((tid & 63) == 0) {
            if (hit) {
                const int lane = __builtin_ctzll(hit);
                const uint64_t ob = (uint64_t)__double_as_longlong(o);
                const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)ob, lane);
                const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(ob >> 32), lane);
                S.rv[tid >> 6] = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
                S.ri[tid >> 6] = ib0 + tid + lane;
            } else {
                S.rv[tid >> 6] = none_v;
                S.ri[tid >> 6] = INT32_MAX;
            }
        }
    }
    // also orders the previous pivot's reads of S.ci / S.xch and its row
    // stores before this pivot's swap
    __syncthreads();
    int im = i;
    double best_order = 0.0;
    if (pivot) {
        // the wavefront minima, in wavefront order (as one thread would):
        // every thread reads the same eight entries
        double bv = S.rv[0];
        int bi = S.ri[0];
#pragma unroll
        for (int w = 1; w < kRegCols / 64; ++w)
            if (pivot_better(pivot, S.rv[w], S.ri[w], bv, bi)) {
                bv = S.rv[w];
                bi = S.ri[w];
            }
        // uniform from here: kept in scalar registers
        bi = __builtin_amdgcn_readfirstlane(bi);
        const uint64_t bb = __double_as_longlong(bv);
        bv = __longlong_as_double((long long)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(bb >> 32)) << 32) |
                                              (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)bb)));
        im = bi;
        best_order = bv;
        // rowmax: a NaN at i keeps i (nothing compares greater)
        if (pivot == kPivotRowMax && (bi == INT32_MAX || isnan(S.dg[a0] - S.dsum[a0]))) im = i;
    }
    const double m_ii = S.dg[im - ib0] - S.dsum[im - ib0];  // read before the swap
    // rank: genz stops at m_ii <= eps0 after the swap (src/cholesky.jl:518-541),
    // rowmax at m_ii <= eps0 before it, eps0 then being the rank tolerance
    // (0 for the reference's m_ii <= 0, :329-350)
    const bool stop = !(m_ii > eps0) && (pivot == kPivotGenz || pivot == kPivotRowMax);
    if (stop && pivot == kPivotRowMax) im = i;
    const int a1 = im - ib0;
    // the swap of the per-index state: the owners of i and im copy their
    // entries to LDS staging now and take each other's after the barrier
    // below (or at once when the block stops here: nobody reads them again in
    // this launch); no registers hold them across the swap
    const bool swp = im != i && (tid == a0 || tid == a1);
    if (swp) {
        const int k = tid == a0 ? 0 : 1;
        S.x_perm[k] = S.perm[tid];
        S.x_dsum[k] = S.dsum[tid];
        S.x_sv[k] = S.sv[tid];
        S.x_dg[k] = S.dg[tid];
        S.x_a[k] = S.a[tid];
        S.x_b[k] = S.b[tid];
    }
    auto take_swapped = [&]() {
        if (swp) {
            const int k = tid == a0 ? 1 : 0;
            S.perm[tid] = S.x_perm[k];
            S.dsum[tid] = S.x_dsum[k];
            S.sv[tid] = S.x_sv[k];
            S.dg[tid] = S.x_dg[k];
            if (pivot != kPivotRowMax) {  // rowmax leaves a and b in place
                S.a[tid] = S.x_a[k];
                S.b[tid] = S.x_b[k];
            }
        }
    };
    if (stop) {
        __syncthreads();  // every thread has read the entries of i and im, and the staging is complete
        take_swapped();
        if (tid == 0) *C.rank = i;
        return false;
    }
    const double uii = sqrt(m_ii);
    double yi = 0.0;
    if (pivot == kPivotGenz) {
        // position i now holds column im's bounds and conditional mean
        const double a_norm = (S.a[a1] - S.sv[a1]) / uii;
        const double b_norm = (S.b[a1] - S.sv[a1]) / uii;
        yi = (exp(-(a_norm * a_norm) / 2) - exp(-(b_norm * b_norm) / 2)) * is2pi / best_order;
    }
    // ---- symmetric swap i <-> im; each window thread keeps the new M[i, c].
    // Every element moves once; the loads of a thread's first rows and
    // columns go out together, so the swap costs one memory round trip.
    double mic = 0.0;
    bool have = false;
    if (im != i) {
        constexpr int U = 2;
        double ri[U], rm[U], ci_[U], cm[U];
        const int rend = min(i, U * kRegCols), cend = min(n, w1 + U * kRegCols);
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int r = tid + k * kRegCols, cc = w1 + tid + k * kRegCols;
            if (r < rend) {  // rows r < i: columns i and im
                ri[k] = MM(r, i);
                rm[k] = MM(r, im);
            }
            if (cc < cend) {  // columns beyond the window: rows i and im
                ci_[k] = MM(i, cc);
                cm[k] = MM(im, cc);
            }
        }
        double x = 0.0, y = 0.0;
        const bool mid = own && c > i && c < im, past = own && c > im;
        if (mid) {  // M[i, c] <-> M[c, im]
            x = MM(i, c);
            y = MM(c, im);
        } else if (past) {  // M[i, c] <-> M[im, c]
            x = MM(i, c);
            y = MM(im, c);
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int r = tid + k * kRegCols, cc = w1 + tid + k * kRegCols;
            if (r < rend) {
                MM(r, i) = rm[k];
                MM(r, im) = ri[k];
            }
            if (cc < cend) {
                MM(i, cc) = cm[k];
                MM(im, cc) = ci_[k];
            }
        }
        if (mid) {
            MM(i, c) = y;
            MM(c, im) = x;
        } else if (past) {
            MM(i, c) = y;
            MM(im, c) = x;
        }
        mic = y;
        have = mid || past;
        for (int r = rend + tid; r < i; r += kRegCols) {  // the rest (i > U kRegCols)
            const double t = MM(r, i);
            MM(r, i) = MM(r, im);
            MM(r, im) = t;
        }
        for (int cc = cend + tid; cc < n; cc += kRegCols) {
            const double t = MM(i, cc);
            MM(i, cc) = MM(im, cc);
            MM(im, cc) = t;
        }
        if (tid == 0) MM(im, im) = MM(i, i);  // MM(i, i) becomes U_ii below
        // the block rows of columns i and im are in registers as well: column
        // i's new rows come from the thread of im, which takes column i's old ones
        if (tid == a1) regs_store<E>(R, ng, S.ci);
        if (tid == a0) regs_store<E>(R, ng, S.xch);
    } else if (tid == a0) {
        regs_store<E>(R, ng, S.ci);
    }
    __syncthreads();
    take_swapped();
    if (im != i && tid == a1) regs_load<E>(R, ng, S.xch);
    if (tid == 0) MM(i, i) = uii;
    // ---- row i of U inside the window (src/cholesky.jl:547-570)
    if (own && c > i) {
        const double acc = regs_dot<E>(R, ng, S.ci, have ? mic : MM(i, c));
        const double u = acc / uii;
        MM(i, c) = u;
        R.cur[E] = u;
        S.dsum[tid] += u * u;
        S.sv[tid] += yi * u;
    }
    return true;
}

template <int PIVOT, int E>
__device__ __forceinline__ bool reg_pivots(const CholArgs& C, RegShared& S, ColRegs& R, int ng, int ib0, int ib1,
                                           int w1, double eps0)
{
    if constexpr (E < 16) {
        if (ib0 + 16 * ng + E >= ib1) return false;  // uniform: the block is done
        if (!reg_pivot<PIVOT, E>(C, S, R, ng, ib0, w1, eps0)) return false;
        return reg_pivots<PIVOT, E + 1>(C, S, R, ng, ib0, ib1, w1, eps0);
    }
    return true;
}

// PIVOT: kPivotNone, kPivotGenz or kPivotRowMax (a template parameter, so that
// each mode's register allocation is its own)
template <int PIVOT>
__global__ __launch_bounds__(kRegCols, 1) void chol_window_reg_kernel(CholArgs C, int w0, int w1, int ib0, int nb)
{
    if (*C.rank < C.n) return;  // an earlier pivot found the matrix rank deficient
    __shared__ RegShared S;
    const double* M = C.M;
    const int64_t lda = C.lda;
    const int tid = threadIdx.x;
    const int ib1 = min(w1, ib0 + nb);
    const int c = ib0 + tid;
    const bool own = c < w1;
    const double eps0 = *C.eps0;
    if (own) {
        S.sv[tid] = C.s_vec[c];
        S.a[tid] = C.a[c];
        S.b[tid] = C.b[c];
        S.perm[tid] = C.perm[c];
        S.dg[tid] = MM(c, c);
        S.dsum[tid] = 0.0;  // d_vec reset at the block start (src/cholesky.jl:494-496)
    }
    ColRegs R;
#pragma unroll
    for (int e = 0; e < 16; ++e) R.g0[e] = R.g1[e] = R.g2[e] = R.cur[e] = 0.0;
    __syncthreads();
    for (int ng = 0; ng < kRegNb / 16; ++ng) {
        if (!reg_pivots<PIVOT, 0>(C, S, R, ng, ib0, ib1, w1, eps0)) break;
        // the completed group moves to its slot
        if (ng == 0) {
#pragma unroll
            for (int e = 0; e < 16; ++e) R.g0[e] = R.cur[e];
        } else if (ng == 1) {
#pragma unroll
            for (int e = 0; e < 16; ++e) R.g1[e] = R.cur[e];
        } else if (ng == 2) {
#pragma unroll
            for (int e = 0; e < 16; ++e) R.g2[e] = R.cur[e];
        }
    }
    if (own) {
        C.s_vec[c] = S.sv[tid];
        C.a[c] = S.a[tid];
        C.b[c] = S.b[tid];
        C.perm[c] = S.perm[tid];
    }
}

hipError_t chol_syrk(const CholArgs& C, int k0, int k1, int r0, int r1, hipStream_t s);
rocblas_handle blas_handle(hipStream_t s);
void* blas_handle_void(hipStream_t s) { return (void*)blas_handle(s); }

hipError_t chol_window(const CholArgs& C, int w0, int w1, int nb, int pivot, hipStream_t s)
{
    for (int ib0 = w0; ib0 < w1; ib0 += nb) {
        const int ib1 = std::min(w1, ib0 + nb);
        if (w1 - ib0 <= kRegCols && nb <= kRegNb) {
            if (pivot == kPivotGenz)
                hipLaunchKernelGGL(chol_window_reg_kernel<kPivotGenz>, dim3(1), dim3(kRegCols), 0, s, C, w0, w1, ib0,
                                   nb);
            else if (pivot == kPivotRowMax)
                hipLaunchKernelGGL(chol_window_reg_kernel<kPivotRowMax>, dim3(1), dim3(kRegCols), 0, s, C, w0, w1,
                                   ib0, nb);
            else
                hipLaunchKernelGGL(chol_window_reg_kernel<kPivotNone>, dim3(1), dim3(kRegCols), 0, s, C, w0, w1, ib0,
                                   nb);
        } else {
            hipLaunchKernelGGL(chol_window_kernel, dim3(1), dim3(kCholThreads), 0, s, C, w0, w1, ib0, nb, pivot);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        // deferred update of the rest of the window by the tiled kernel
        if (ib1 < w1) {
            e = chol_syrk(C, ib0, ib1, ib1, w1, s);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

// One rocBLAS handle per host thread and device, created on first use and
// kept until the process ends; its stream is set before every call.
// (Declared for the other translation units as blas_handle_void.)
rocblas_handle blas_handle(hipStream_t s)
{
    constexpr int kMaxDev = 64;
    thread_local rocblas_handle handles[kMaxDev] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) return nullptr;
    if (!handles[dev] && rocblas_create_handle(&handles[dev]) != rocblas_status_success) {
        handles[dev] = nullptr;
        return nullptr;
    }
    if (rocblas_set_stream(handles[dev], s) != rocblas_status_success) return nullptr;
    return handles[dev];
}

// trailing update M[x, y] -= sum_{l in [k0, k1)} U[l, x] U[l, y] for r0 <= x <= y < r1:
// per element one FMA chain in ascending l from zero, then the subtraction.
// 64 x 64 tiles, 4 x 4 per thread; stages of 32 l in two LDS buffers, the
// next stage's loads held in registers while the current one is multiplied,
// so a stage costs one barrier and its load latency hides behind the products
constexpr int kSyTile = 64;
constexpr int kSyK = 32;
constexpr int kSyLd = kSyK * kSyTile / 256;  // elements of each operand a thread stages
__global__ __launch_bounds__(256) void chol_syrk_kernel(double* M, int64_t lda, int n, int k0, int k1,
                                                        int r0, int r1, const int32_t* rank)
{
    if (*rank < n) return;
    const int tx = blockIdx.x, ty = blockIdx.y;
    if (tx > ty) return;
    const int x0 = r0 + tx * kSyTile, y0 = r0 + ty * kSyTile;
    __shared__ double Ax[2][kSyK][kSyTile + 1], Ay[2][kSyK][kSyTile + 1];
    const int t = threadIdx.x;
    const int tr = t % 16, tc = t / 16;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    // staging: element e = t + 256 q is (l0 + e mod kSyK, column e / kSyK)
    double ra[kSyLd], rb[kSyLd];
    auto fetch = [&](int l0) {
#pragma unroll
        for (int q = 0; q < kSyLd; ++q) {
            const int e = t + 256 * q;
            const int l = l0 + e % kSyK, cc = e / kSyK;
            ra[q] = (l < k1 && x0 + cc < r1) ? MM(l, x0 + cc) : 0.0;
            rb[q] = (l < k1 && y0 + cc < r1) ? MM(l, y0 + cc) : 0.0;
        }
    };
    auto put = [&](int buf) {
#pragma unroll
        for (int q = 0; q < kSyLd; ++q) {
            const int e = t + 256 * q;
            Ax[buf][e % kSyK][e / kSyK] = ra[q];
            Ay[buf][e % kSyK][e / kSyK] = rb[q];
        }
    };
    fetch(k0);
    put(0);
    __syncthreads();
    int buf = 0;
    for (int l0 = k0; l0 < k1; l0 += kSyK) {
        const bool more = l0 + kSyK < k1;  // uniform
        if (more) fetch(l0 + kSyK);
#pragma unroll 4
        for (int ll = 0; ll < kSyK; ++ll) {
            double xa[4], ya[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) xa[i] = Ax[buf][ll][tr * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) ya[j] = Ay[buf][ll][tc * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(xa[i], ya[j], acc[i][j]);
        }
        // the other buffer was last read before the previous stage's barrier
        if (more) put(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int x = x0 + tr * 4 + i, y = y0 + tc * 4 + j;
            if (x <= y && y < r1) MM(x, y) -= acc[i][j];
        }
}

hipError_t chol_syrk(const CholArgs& C, int k0, int k1, int r0, int r1, hipStream_t s)
{
    const int rest = r1 - r0;
    if (rest <= 0 || k1 <= k0) return hipSuccess;
    const int nt = (rest + kSyTile - 1) / kSyTile;
    hipLaunchKernelGGL(chol_syrk_kernel, dim3(nt, nt), dim3(256), 0, s, C.M, C.lda, C.n, k0, k1, r0, r1,
                       C.rank);
    return hipGetLastError();
}

// Rows [r0, r1) of U in columns [c0, c1), U_rr^T X = M[r0:r1, c0:c1] with
// the factored diagonal block U_rr = M[r0:r1, r0:r1] (a plain triangular
// solve, left / upper / transposed / non-unit), then the trailing update of
// [c0, c1) by those rows: the ldiv! and mul! of one level of the reference's
// blocking (src/cholesky.jl:589-624; r1 = c0, c1 the parent block's end).
hipError_t chol_trsm_syrk(const CholArgs& C, int r0, int r1, int c0, int c1, hipStream_t s)
{
    const int rest = c1 - c0;
    if (rest <= 0 || r1 <= r0) return hipSuccess;
    rocblas_handle h = blas_handle(s);
    if (!h) return hipErrorNotInitialized;
    const double one = 1.0;
    double* M = C.M;
    const int64_t lda = C.lda;
    const rocblas_status st =
        rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                      rocblas_diagonal_non_unit, r1 - r0, rest, &one, &MM(r0, r0), (rocblas_int)lda,
                      &MM(r0, c0), (rocblas_int)lda);
    if (st != rocblas_status_success) return hipErrorLaunchFailure;
    return chol_syrk(C, r0, r1, c0, c1, s);
}

__global__ void chol_zero_lower_kernel(double* M, int64_t lda, int n)
{
    const int j = blockIdx.y;
    for (int i = j + 1 + blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        MM(i, j) = 0.0;
}

hipError_t chol_finish(const CholArgs& C, hipStream_t s)
{
    dim3 g((C.n + 255) / 256, C.n);
    hipLaunchKernelGGL(chol_zero_lower_kernel, g, dim3(256), 0, s, C.M, C.lda, C.n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// adapt_low_rank!  src/cholesky.jl:20-123, one workgroup.  Columns rank .. n-1
// of the factor (the dependent variables) are taken in turn: the level i_j of
// column j is its last row below rank with |M[i, j]| > eps1 (a parallel max
// over the rows); a column without one moves to the end (swap with the last
// active column, k_0 += 1); otherwise rows 0 .. i_j are scaled by 1/M[i_j, j]
// (a and b too, exchanged when that entry is negative) and the column is
// inserted after the block of its level, end[i_j] + 1, by rotating the columns
// in between right by one through the scratch buffer (two element-parallel
// passes).  end[i] is the 0-based last column of level i's block (the
// reference's k_js minus one).  Every branch depends on values broadcast
// through LDS, so the control flow is uniform.
// ---------------------------------------------------------------------------
constexpr int kAlrThreads = 1024;

__global__ __launch_bounds__(kAlrThreads) void adapt_low_rank_kernel(double* __restrict__ M, int64_t lda, int n,
                                                                    double* __restrict__ a, double* __restrict__ b,
                                                                    int32_t* __restrict__ perm, int rank,
                                                                    double eps1, int32_t* __restrict__ end,
                                                                    double* __restrict__ tmp,
                                                                    int32_t* __restrict__ k0_out)
{
    __shared__ int s_lvl[kAlrThreads / 64];
    __shared__ int s_ij, s_j, s_nact, s_k0;
    __shared__ double s_mij;
    const int tid = threadIdx.x;
    for (int i = tid; i < rank; i += kAlrThreads) end[i] = i;
    if (tid == 0) {
        s_j = rank;
        s_nact = n;
        s_k0 = 0;
    }
    __syncthreads();
    for (;;) {
        const int j = s_j, nact = s_nact;
        if (j >= nact) break;  // uniform
        // 1. the level: last row i < rank with |M[i, j]| > eps1 (-1: none)
        int lv = -1;
        for (int i = tid; i < rank; i += kAlrThreads)
            if (fabs(MM(i, j)) > eps1) lv = i;
        lv = wave_max_int(lv);
        if ((tid & 63) == 0) s_lvl[tid >> 6] = lv;
        __syncthreads();
        if (tid == 0) {
            int m = -1;
            for (int w = 0; w < kAlrThreads / 64; ++w) m = max(m, s_lvl[w]);
            s_ij = m;
            // the pivot entry, read once before the scaling below rewrites it
            s_mij = m >= 0 ? MM(m, j) : 0.0;
        }
        __syncthreads();
        const int ij = s_ij;
        if (ij < 0) {
            // a zero column: to the end of the active range
            const int last = nact - 1;
            if (j != last) {
                for (int r = tid; r < n; r += kAlrThreads) {
                    const double t = MM(r, last);
                    MM(r, last) = MM(r, j);
                    MM(r, j) = t;
                }
                if (tid == 0) {
                    double t = a[j]; a[j] = a[last]; a[last] = t;
                    t = b[j]; b[j] = b[last]; b[last] = t;
                    const int32_t p = perm[j]; perm[j] = perm[last]; perm[last] = p;
                }
            }
            if (tid == 0) {
                s_k0 += 1;
                s_nact = last;  // j stays: the column swapped in is taken next
            }
            __syncthreads();
            continue;
        }
        // 2. scale the column up to its level, and the bounds
        const double mij = s_mij;
        const double inv = 1.0 / mij;
        for (int r = tid; r <= ij; r += kAlrThreads) MM(r, j) *= inv;
        if (tid == 0) {
            a[j] *= inv;
            b[j] *= inv;
            if (mij < 0.0) {
                const double t = a[j];
                a[j] = b[j];
                b[j] = t;
            }
        }
        const int pos = end[ij] + 1;
        __syncthreads();
        // 3. insert column j at pos: columns pos .. j-1 move right by one
        if (j != pos) {
            const int ncol = j - pos + 1;
            const int64_t cnt = (int64_t)ncol * n;
            for (int64_t e = tid; e < cnt; e += kAlrThreads) {
                const int c = pos + (int)(e / n), r = (int)(e % n);
                tmp[e] = MM(r, c);
            }
            __syncthreads();
            for (int64_t e = tid; e < cnt; e += kAlrThreads) {
                const int c = pos + (int)(e / n), r = (int)(e % n);
                MM(r, c == j ? pos : c + 1) = tmp[e];
            }
            if (tid == 0) {
                const double va = a[j], vb = b[j];
                const int32_t vp = perm[j];
                for (int c = j; c > pos; --c) {
                    a[c] = a[c - 1];
                    b[c] = b[c - 1];
                    perm[c] = perm[c - 1];
                }
                a[pos] = va;
                b[pos] = vb;
                perm[pos] = vp;
            }
        }
        // 4. the blocks of levels ij .. rank-1 end one column later
        for (int i = ij + tid; i < rank; i += kAlrThreads) end[i] += 1;
        if (tid == 0) s_j = j + 1;
        __syncthreads();
    }
    if (tid == 0) *k0_out = s_k0;
}

hipError_t adapt_low_rank(double* M, int64_t lda, int n, double* a, double* b, int32_t* perm, int rank, double eps1,
                          int32_t* end, double* tmp, int32_t* k0, hipStream_t s)
{
    hipLaunchKernelGGL(adapt_low_rank_kernel, dim3(1), dim3(kAlrThreads), 0, s, M, lda, n, a, b, perm, rank, eps1,
                       end, tmp, k0);
    return hipGetLastError();
}

#undef MM
}  // namespace dbg

// ---------------------------------------------------------------------------
// Batched constructor: one workgroup per problem of a batch of uneven
// dimension (d <= kBatchMaxDim, a single pivot window).  It runs the whole
// QMCData constructor for its problem on the device: standardize, pivoted
// factor, truncate_matrixU!, and the layouts read by the integration kernel
// (packed rows, column-major copy, bounds, 1/(U_jj sqrt 2), first interval).
// ---------------------------------------------------------------------------
namespace dbg {

template <bool MIXED>
__global__ __launch_bounds__(256) void batch_factor_kernel(BatchArgs B,
                                                          typename std::conditional<MIXED, float, double>::type* __restrict__ Uk,
                                                          typename std::conditional<MIXED, float, double>::type* __restrict__ Ucol)
{
    using T = typename std::conditional<MIXED, float, double>::type;
    __shared__ double s_red[4];
    const int p = blockIdx.x;
    const BatchProblem P = B.probs[p];
    const int d = P.d;
    const int tid = threadIdx.x;
    double* M = B.M + P.sig_off;
    double* a = B.a + P.vec_off;
    double* b = B.b + P.vec_off;
    double* scale = B.scale + P.vec_off;
    double* sv = B.svec + P.vec_off;
    int32_t* perm = B.perm + P.vec_off;
    const int64_t lda = d;
    for (int64_t e = tid; e < (int64_t)d * d; e += 256) M[e] = B.sig_in[P.sig_off + e];
    double mx = 0.0;
    for (int j = tid; j < d; j += 256) {
        a[j] = B.a_in[P.vec_off + j];
        b[j] = B.b_in[P.vec_off + j];
        sv[j] = 0.0;
        perm[j] = j;
        // the diagonal from the input: the copy into M above is made by other
        // threads and is not complete before the barrier below
        const double v = B.sig_in[P.sig_off + (int64_t)j * (d + 1)];
        const double av = fabs(v);
        mx = (isnan(av) || av > mx) ? av : mx;
        scale[j] = v > 0 ? sqrt(v) : 1.0;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(mx, off, 64);
        mx = (isnan(o) || o > mx) ? o : mx;
    }
    if ((tid & 63) == 0) s_red[tid >> 6] = mx;
    if (tid == 0) B.rank[p] = d;
    __syncthreads();
    double m4 = s_red[0];
    for (int w = 1; w < 4; ++w) m4 = (isnan(s_red[w]) || s_red[w] > m4) ? s_red[w] : m4;
    const double eps0 = DBL_EPSILON * m4;
    // standardize (src/cholesky.jl:458-473), upper triangle
    for (int64_t e = tid; e < (int64_t)d * d; e += 256) {
        const int i = (int)(e % d), j = (int)(e / d);
        if (i < j) M[e] = M[e] * ((1.0 / scale[j]) / scale[i]);
        else if (i == j) M[e] = 1.0;
    }
    __syncthreads();
    for (int j = tid; j < d; j += 256) {
        const double inv = 1.0 / scale[j];
        a[j] *= inv;
        b[j] *= inv;
    }
    __syncthreads();
    CholArgs C{M, lda, d, a, b, scale, perm, sv, nullptr, B.rank + p};
    chol_window_body<256, kBatchMaxDim>(C, 0, d, 0, d, max(1, min(B.bs1, 64)), 1, eps0, true);
    const bool ok = B.rank[p] == d;
    // zero the strict lower triangle, truncate_matrixU! (src/qmc.jl:133-142)
    const double min_val = MIXED ? (double)(FLT_EPSILON / (float)(d * 9)) : DBL_EPSILON / (double)(d * 9);
    for (int64_t e = tid; e < (int64_t)d * d; e += 256) {
        const int i = (int)(e % d), j = (int)(e / d);
        double v = (i > j) ? 0.0 : M[e];
        if (fabs(v) < min_val) v = 0.0;
        M[e] = v;
    }
    __syncthreads();
    T* uk = Uk + P.uk_off;
    T* uc = Ucol + P.ucol_off;
    for (int64_t e = tid; e < (int64_t)d * d; e += 256) uc[e] = (T)M[e];
    for (int64_t e = tid; e < (int64_t)max(d - 1, 0) * P.ldu; e += 256) {
        const int l = (int)(e / P.ldu), j = (int)(e % P.ldu);
        uk[e] = (T)((j > l && j < d) ? M[(size_t)l + (size_t)j * lda] : 0.0);
    }
    T* ak = (T*)B.ak + P.vec_off;
    T* bk = (T*)B.bk + P.vec_off;
    T* ik = (T*)B.invk + P.vec_off;
    for (int j = tid; j < d; j += 256) {
        const double ujj = M[(size_t)j * (d + 1)];
        ak[j] = (T)a[j];
        bk[j] = (T)b[j];
        ik[j] = (T)(1.0 / (ujj * 1.4142135623730951));
    }
    if (tid == 0) {
        double* f = B.first + 8 * (int64_t)p;
        const double u11 = M[0];
        const double arg_a = a[0] / (u11 * 1.4142135623730951);
        const double arg_b = b[0] / (u11 * 1.4142135623730951);
        f[0] = 0.5 * (1.0 + erf(arg_a));
        f[1] = 0.5 * (erf(arg_b) - erf(arg_a));
        const double ea = 0.5 * erfc(fabs(arg_a));
        const double eb = 0.5 * erfc(fabs(arg_b));
        if (arg_b <= 0.0) {
            f[4] = eb - ea;
            f[5] = 0;
        } else if (arg_a >= 0.0) {
            f[4] = ea - eb;
            f[5] = 1;
        } else {
            f[4] = 0.5 * (erf(arg_b) - erf(arg_a));
            f[5] = 2;
        }
        f[2] = ea;
        f[3] = eb;
        f[6] = ok ? 1.0 : 0.0;
    }
}

hipError_t batch_factor(const BatchArgs& B, bool mixed, void* Uk, void* Ucol, hipStream_t s)
{
    if (mixed)
        hipLaunchKernelGGL(batch_factor_kernel<true>, dim3(B.K), dim3(256), 0, s, B, (float*)Uk, (float*)Ucol);
    else
        hipLaunchKernelGGL(batch_factor_kernel<false>, dim3(B.K), dim3(256), 0, s, B, (double*)Uk, (double*)Ucol);
    return hipGetLastError();
}

}  // namespace dbg
