// Helper kernels of the adjoint (gradient of log P; no reference
// counterpart).  The sweep itself runs in the matrix-core integration kernel
// (record and reverse modes); these kernels prepare its inputs and turn its
// outputs into the gradient with respect to the original a, b and Sigma:
//
//   pack_flip      the flipped factor Uf[g1, g2] = U[n-1-g2, n-1-g1] in the
//                  packed-row and column-major layouts of the dense kernel,
//                  so the reverse recursion ybar_l = sum_{r>l} U[l, r] sbar_r
//                  is the forward contraction in flipped index order;
//   grad_norm      E = max exponent of the tile sums, 1 / sum(p 2^-E) and
//                  log P of the recorded forward (one chunk of samples);
//   grad_fold_*    the weights with which a chunk's gradient joins those of
//                  the chunks before it (each chunk's share of the sum);
//   ubar_product   Ubar[l, g+1] = sum_s y_l sbar_{g+1} (l <= g), split-K on
//                  the matrix cores with FP64 sums of the chunk partials;
//   tril_half      Phi(X): lower triangle, halved diagonal (Cholesky adjoint);
//   sigma_map      C_bar -> G (symmetrized, unpermuted, unstandardized) and
//                  the gradients of the original bounds;
//   sigma_diag     the diagonal terms from the standardization and the bounds.
#include <algorithm>

#include "dbg_device.h"

namespace dbg {

template <typename T>
__global__ void pack_flip_kernel(const double* __restrict__ U, int n, T* __restrict__ Ufk, int64_t ldu,
                                 T* __restrict__ Ufc)
{
    const int g1 = blockIdx.y;  // 0 .. n-1
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ldu; g += (int64_t)gridDim.x * blockDim.x) {
        const bool in = g < n;
        const double v = in ? U[(size_t)(n - 1 - g) + (size_t)(n - 1 - g1) * n] : 0.0;
        if (g1 < n - 1) Ufk[(size_t)g1 * ldu + g] = (T)((g > g1 && in) ? v : 0.0);
        if (in) Ufc[(size_t)g * n + g1] = (T)(g1 <= g ? v : 0.0);
    }
}

hipError_t pack_flip(const double* U, int n, bool mixed, void* Ufk, int64_t ldu, void* Ufc, hipStream_t s)
{
    dim3 g((unsigned)std::min<int64_t>((ldu + 255) / 256, 64), n), b(256);
    if (mixed)
        hipLaunchKernelGGL(pack_flip_kernel<float>, g, b, 0, s, U, n, (float*)Ufk, ldu, (float*)Ufc);
    else
        hipLaunchKernelGGL(pack_flip_kernel<double>, g, b, 0, s, U, n, (double*)Ufk, ldu, (double*)Ufc);
    return hipGetLastError();
}

// norm[0] = E, norm[1] = 1 / S with S = sum_t part_m[t] 2^(part_e[t] - E);
// norm[2] = log P = log(S / total) + E log 2, norm[3] = S.  One block.
__global__ __launch_bounds__(256) void grad_norm_kernel(const double* __restrict__ pm, const int* __restrict__ pe,
                                                        int64_t n_tiles, double total, double* norm)
{
    __shared__ int s_e[4];
    __shared__ double s_v[256], s_c[256];
    const int tid = threadIdx.x;
    int e = INT32_MIN;
    for (int64_t t = tid; t < n_tiles; t += 256)
        if (pm[t] != 0.0) e = max(e, pe[t]);
    e = wave_max_int(e);
    if ((tid & 63) == 0) s_e[tid >> 6] = e;
    __syncthreads();
    int E = max(max(s_e[0], s_e[1]), max(s_e[2], s_e[3]));
    if (E == INT32_MIN) E = 0;
    double sum = 0.0, comp = 0.0;
    for (int64_t t = tid; t < n_tiles; t += 256) {
        const double v = pm[t] != 0.0 ? ldexp(pm[t], pe[t] - E) : 0.0;
        const double tt = sum + v;
        comp += (fabs(sum) >= fabs(v)) ? (sum - tt) + v : (v - tt) + sum;
        sum = tt;
    }
    s_v[tid] = sum;
    s_c[tid] = comp;
    __syncthreads();
    if (tid == 0) {
        double S = 0.0, C = 0.0;
        for (int q = 0; q < 256; ++q) {
            const double v = s_v[q];
            const double tt = S + v;
            C += (fabs(S) >= fabs(v)) ? (S - tt) + v : (v - tt) + S;
            S = tt;
            C += s_c[q];
        }
        S += C;
        norm[0] = (double)E;
        norm[1] = S > 0.0 ? 1.0 / S : 0.0;
        norm[2] = S > 0.0 ? log(S / total) + E * 0.69314718055994530942 : -INFINITY;
        norm[3] = S;
    }
}

hipError_t grad_norm(const double* pm, const int* pe, int64_t n_tiles, double total, double* norm, hipStream_t s)
{
    hipLaunchKernelGGL(grad_norm_kernel, dim3(1), dim3(256), 0, s, pm, pe, n_tiles, total, norm);
    return hipGetLastError();
}

// The gradient of log P over all samples is the weighted mean of the chunks'
// gradients (each normalised by its own sum), with weights W_c / W, W_c =
// S_c 2^E_c: with the running sums normalised by W_r of the chunks before,
// the new running sum is alpha (old) + beta (chunk), alpha = W_r / (W_r + W_c),
// beta = W_c / (W_r + W_c).  run = {E_r, S_r} (S_r in units of 2^E_r).
__global__ void grad_fold_norm_kernel(double* norm, double* run, double* fold, int first, int last, double total)
{
    if (threadIdx.x != 0) return;
    const double Ec = norm[0], Sc = norm[3];
    double Er = first ? Ec : run[0], Sr = first ? 0.0 : run[1];
    double alpha = 0.0, beta = 0.0;
    if (Sr <= 0.0) {
        Er = Ec;
        Sr = Sc;
        beta = Sc > 0.0 ? 1.0 : 0.0;
    } else if (Sc > 0.0) {
        const double E = fmax(Er, Ec);
        const double u = ldexp(Sr, (int)(Er - E)), v = ldexp(Sc, (int)(Ec - E));
        Er = E;
        Sr = u + v;
        alpha = u / Sr;
        beta = v / Sr;
    } else {
        alpha = 1.0;
    }
    run[0] = Er;
    run[1] = Sr;
    fold[0] = alpha;
    fold[1] = beta;
    if (last) norm[2] = Sr > 0.0 ? log(Sr / total) + Er * 0.69314718055994530942 : -INFINITY;
}

hipError_t grad_fold_norm(double* norm, double* run, double* fold, int first, int last, double total,
                          hipStream_t s)
{
    hipLaunchKernelGGL(grad_fold_norm_kernel, dim3(1), dim3(64), 0, s, norm, run, fold, first, last, total);
    return hipGetLastError();
}

__global__ void grad_fold_acc_kernel(double* __restrict__ acc, const double* __restrict__ cacc, int len,
                                     const double* __restrict__ fold)
{
    const double alpha = fold[0], beta = fold[1];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < len; i += gridDim.x * blockDim.x)
        acc[i] = alpha * acc[i] + beta * cacc[i];
}

hipError_t grad_fold_acc(double* acc, const double* cacc, int len, const double* fold, hipStream_t s)
{
    hipLaunchKernelGGL(grad_fold_acc_kernel, dim3((len + 255) / 256), dim3(256), 0, s, acc, cacc, len, fold);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Ubar[l, g + 1] = sum_s y_l(s) sbar_{g+1}(s): a product of two row-major
// (row, sample) arrays over a very long sample axis and a short output
// (nd x nd, upper triangle only).  Split-K on the matrix cores: workgroup
// (tile pair, chunk) computes a 128 x 128 output tile over one chunk of
// samples (each wave 64 x 64, 4 x 4 blocks of 16 x 16); lane (r, q) loads
// four consecutive samples of one row with one vector load and feeds them to
// four K = 4 instructions (sample s + 4q + c is the k index q of the c-th),
// so both operands stream straight from memory in 64-byte row segments.
// The chunk partials go to `part` and are summed in FP64 by ubar_reduce.
// ---------------------------------------------------------------------------
template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
    typedef float V __attribute__((ext_vector_type(4)));
};
template <>
struct Vec4<double> {
    typedef double V __attribute__((ext_vector_type(4)));
};

constexpr int kUbTile = 128;

template <typename T>
__global__ __launch_bounds__(256) void ubar_gemm_kernel(const T* __restrict__ Y, const T* __restrict__ S,
                                                        int64_t rs, int nd, const int2* __restrict__ pairs,
                                                        int n_pairs, int64_t kchunk, int64_t kmax,
                                                        T* __restrict__ part)
{
    using V = typename Mfma<T>::V;
    using L4 = typename Vec4<T>::V;
    const int pair = blockIdx.x, kc = blockIdx.y;
    const int2 tp = pairs[pair];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    const int wr = wv & 1, wc = wv >> 1;
    const int l0 = tp.x * kUbTile + wr * 64, g0 = tp.y * kUbTile + wc * 64;
    const int64_t k0 = (int64_t)kc * kchunk;
    const int64_t k1 = min(kmax, k0 + kchunk);
    const T* __restrict__ pa[4];
    const T* __restrict__ pb[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int l = min(l0 + 16 * q + lr, nd - 1);  // rows past nd read row nd-1; discarded
        const int g = min(g0 + 16 * q + lr, nd - 1);
        pa[q] = Y + (size_t)l * rs + 4 * lq;
        pb[q] = S + (size_t)g * rs + 4 * lq;
    }
    V acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = V{0, 0, 0, 0};
    L4 a[2][4], b[2][4];
    if (k0 < k1) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            a[0][q] = *(const L4*)(pa[q] + k0);
            b[0][q] = *(const L4*)(pb[q] + k0);
        }
    }
    for (int64_t s = k0; s < k1; s += 32) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int64_t sn = min(s + 16 * (h + 1), k1 - 16);  // next step (re-reads the last at the end)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[h ^ 1][q] = *(const L4*)(pa[q] + sn);
                b[h ^ 1][q] = *(const L4*)(pb[q] + sn);
            }
            if (s + 16 * h < k1) {
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::mac(a[h][i][c], b[h][j][c], acc[i][j]);
            }
        }
    }
    T* __restrict__ out = part + ((size_t)kc * n_pairs + pair) * (kUbTile * kUbTile);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                out[(size_t)(wr * 64 + 16 * i + Mfma<T>::row(lq, r)) * kUbTile + wc * 64 + 16 * j + lr] = acc[i][j][r];
}

// FP32 form of the same product with both operand tiles staged in LDS: per
// stage of kUbK samples each thread loads 16 samples of one A row and one B
// row (four 16-byte loads each) for the next stage while the four waves run
// the current one from LDS, so every operand byte is fetched once per
// workgroup instead of once per wave.  Chunk lengths are multiples of kUbK.
constexpr int kUbK = 32;
constexpr int kUbKS = kUbK + 4;  // LDS row stride (floats)
constexpr int kUbE = kUbK / 8;   // 16-byte loads per thread and operand per stage
constexpr int kUbOcc = 2;        // workgroups per CU the registers are budgeted for

__global__ __launch_bounds__(256, kUbOcc) void ubar_gemm_lds_kernel(const float* __restrict__ Y,
                                                               const float* __restrict__ S, int64_t rs, int nd,
                                                               const int2* __restrict__ pairs, int n_pairs,
                                                               int64_t kchunk, int64_t kmax,
                                                               float* __restrict__ part)
{
    using V = Mfma<float>::V;
    using L4 = Vec4<float>::V;
    __shared__ float s_a[2][kUbTile * kUbKS];
    __shared__ float s_b[2][kUbTile * kUbKS];
    const int pair = blockIdx.x, kc = blockIdx.y;
    const int2 tp = pairs[pair];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    const int wr = wv & 1, wc = wv >> 1;
    // on a diagonal tile pair the wave of rows 64-127 x columns 0-63 lies below
    // the diagonal (l > g): it stages operands but runs no products, and its
    // part of the partial tile is never read
    const bool idle = tp.x == tp.y && wr > wc;
    const int64_t k0 = (int64_t)kc * kchunk;
    const int64_t k1 = min(kmax, k0 + kchunk);
    // staging: thread -> tile row tid / 2, samples 16 (tid % 2) .. + 15 of a stage
    const int srow = tid >> 1, shalf = (tid & 1) * (kUbK / 2);
    const int la = min(tp.x * kUbTile + srow, nd - 1);  // rows past nd read row nd-1; discarded
    const int lb = min(tp.y * kUbTile + srow, nd - 1);
    const float* __restrict__ ga = Y + (size_t)la * rs + shalf;
    const float* __restrict__ gb = S + (size_t)lb * rs + shalf;
    float* __restrict__ wa = &s_a[0][srow * kUbKS + shalf];
    float* __restrict__ wb = &s_b[0][srow * kUbKS + shalf];
    V acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = V{0, 0, 0, 0};
    const int nst = (int)((k1 - k0) / kUbK);
    L4 sa[kUbE], sb[kUbE];
    if (nst > 0) {
#pragma unroll
        for (int e = 0; e < kUbE; ++e) {
            sa[e] = *(const L4*)(ga + k0 + 4 * e);
            sb[e] = *(const L4*)(gb + k0 + 4 * e);
        }
#pragma unroll
        for (int e = 0; e < kUbE; ++e) {
            *(L4*)(wa + 4 * e) = sa[e];
            *(L4*)(wb + 4 * e) = sb[e];
        }
        __syncthreads();
    }
    for (int st = 0; st < nst; ++st) {
        const int buf = st & 1;
        if (st + 1 < nst) {  // uniform
            const int64_t kn = k0 + (int64_t)(st + 1) * kUbK;
#pragma unroll
            for (int e = 0; e < kUbE; ++e) {
                sa[e] = *(const L4*)(ga + kn + 4 * e);
                sb[e] = *(const L4*)(gb + kn + 4 * e);
            }
        }
        const float* __restrict__ ra = &s_a[buf][(wr * 64 + lr) * kUbKS + 4 * lq];
        const float* __restrict__ rb = &s_b[buf][(wc * 64 + lr) * kUbKS + 4 * lq];
#pragma unroll
        for (int h = 0; h < kUbK / 16; ++h) {
            if (idle) break;  // uniform
            L4 a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[q] = *(const L4*)(ra + 16 * q * kUbKS + 16 * h);
                b[q] = *(const L4*)(rb + 16 * q * kUbKS + 16 * h);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<float>::mac(a[i][c], b[j][c], acc[i][j]);
        }
        if (st + 1 < nst) {
            const int nb = (st + 1) & 1;
            float* __restrict__ wan = &s_a[nb][srow * kUbKS + shalf];
            float* __restrict__ wbn = &s_b[nb][srow * kUbKS + shalf];
#pragma unroll
            for (int e = 0; e < kUbE; ++e) {
                *(L4*)(wan + 4 * e) = sa[e];
                *(L4*)(wbn + 4 * e) = sb[e];
            }
        }
        __syncthreads();
    }
    if (idle) return;
    float* __restrict__ out = part + ((size_t)kc * n_pairs + pair) * (kUbTile * kUbTile);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                out[(size_t)(wr * 64 + 16 * i + Mfma<float>::row(lq, r)) * kUbTile + wc * 64 + 16 * j + lr] =
                    acc[i][j][r];
}

// Mixed form on the FP16 matrix cores: per stage of 32 samples (one
// instruction's K) both operands are split into high and low FP16 parts as in
// the split-FP16 integration kernel (y' = 2^6 y; three products per pair,
// FP32 accumulation).  The sbar values carry the sample weights and span far
// more than the FP16 range, so each staged row of S (32 samples of one index
// g) gets its own power-of-two scale that brings its largest magnitude into
// [2^9, 2^10); the three products of a stage start from zero and the stage's
// result is added to the FP32 tile sums with each column's scale.  The
// staging threads convert the FP32 records; the tiles go to LDS as [row][32
// high | 32 low] halves with the 16-byte granules of row r permuted by r mod 8
// (conflict-free operand reads).
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
constexpr int kUsRow = 64;  // halves per staged row (32 samples, high and low)
constexpr int kUbTile16 = 256;  // output tile of the split form

__device__ __forceinline__ void ub_split(float v, _Float16& hi, _Float16& lo)
{
    v = fabsf(v) > 65504.0f ? copysignf(65504.0f, v) : v;
    hi = (_Float16)v;
    lo = (_Float16)(v - (float)hi);
}

// TT: output tile (128: four wavefronts of 64 x 64, two workgroups per CU;
// 256: eight wavefronts of 64 x 128, one workgroup per CU, which reads each
// operand row for half as many tile pairs)
template <int TT>
struct UbSplitCfg {
    static constexpr int NT = 2 * TT;           // threads: two per staged row
    static constexpr int WC = TT == 256 ? 128 : 64;  // columns per wavefront
    static constexpr int NJ = WC / 16;
    static constexpr int NWC = TT / WC;         // wavefront columns
    static constexpr int OCC = TT == 256 ? 1 : 2;
};

template <int TT>
__global__ __launch_bounds__(UbSplitCfg<TT>::NT, UbSplitCfg<TT>::OCC) void ubar_gemm_split_kernel(
    const float* __restrict__ Y, const float* __restrict__ S, int64_t rs, int nd, const int2* __restrict__ pairs,
    int n_pairs, int64_t kchunk, int64_t kmax, float* __restrict__ part)
{
    using V = Mfma<float>::V;
    using L4 = Vec4<float>::V;
    using Cfg = UbSplitCfg<TT>;
    constexpr int NJ = Cfg::NJ;
    __shared__ __attribute__((aligned(16))) _Float16 s_a[2][TT * kUsRow];
    __shared__ __attribute__((aligned(16))) _Float16 s_b[2][TT * kUsRow];
    __shared__ float s_sc[2][TT];  // column scale of each staged S row (2^-10+e, with y's 2^-6)
    // XCD-aware order: workgroups are dealt to the eight XCDs round-robin by
    // dispatch index, so consecutive logical indices (the pairs of one chunk,
    // which read the same sample range of a few row tiles) are given to one
    // XCD and share its L2 (when the grid divides by 8; else dispatch order).
    const int64_t did = blockIdx.x + (int64_t)gridDim.x * blockIdx.y;
    const int64_t total = (int64_t)gridDim.x * gridDim.y;
    const int64_t lid = (total & 7) ? did : (did & 7) * (total >> 3) + (did >> 3);
    const int pair = (int)(lid % n_pairs), kc = (int)(lid / n_pairs);
    const int2 tp = pairs[pair];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    const int wr = wv / Cfg::NWC, wc = wv % Cfg::NWC;
    // on a diagonal tile pair, wavefronts whose rows lie below all their columns
    // stage operands but run no products
    const bool idle = tp.x == tp.y && wr * 64 > wc * Cfg::WC + Cfg::WC - 1;
    const int64_t k0 = (int64_t)kc * kchunk;
    const int64_t k1 = min(kmax, k0 + kchunk);
    // staging: thread -> tile row tid / 2, samples 16 (tid % 2) .. + 15 of a stage
    const int srow = tid >> 1, shalf = tid & 1;
    const int la = min(tp.x * TT + srow, nd - 1);  // rows past nd read row nd-1; discarded
    const int lb = min(tp.y * TT + srow, nd - 1);
    const float* __restrict__ ga = Y + (size_t)la * rs + 16 * shalf;
    const float* __restrict__ gb = S + (size_t)lb * rs + 16 * shalf;
    // granule positions of this thread's staged samples: high granules 2h, 2h+1, low 4+2h, 5+2h
    const int sw = srow & 7;
    const int phA = ((2 * shalf) ^ sw) * 8, phB = ((2 * shalf + 1) ^ sw) * 8;
    const int plA = ((4 + 2 * shalf) ^ sw) * 8, plB = ((5 + 2 * shalf) ^ sw) * 8;
    // operand reads: row 16 i + lr (mod 8 = lr mod 8), granules lq / 4 + lq
    const int rh = (lq ^ (lr & 7)) * 8, rl = ((4 + lq) ^ (lr & 7)) * 8;
    V tot[4][NJ];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) tot[i][j] = V{0, 0, 0, 0};
    const int nst = (int)((k1 - k0) / kUbK);
    L4 va[4], vb[4];
    auto fetch = [&](int64_t kn) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            va[e] = *(const L4*)(ga + kn + 4 * e);
            vb[e] = *(const L4*)(gb + kn + 4 * e);
        }
    };
    auto stage = [&](int buf) {
        half8 h0, h1, l0, l1;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            _Float16 hi, lo;
            ub_split(va[e >> 2][e & 3] * 64.0f, hi, lo);
            if (e < 8) { h0[e] = hi; l0[e] = lo; } else { h1[e - 8] = hi; l1[e - 8] = lo; }
        }
        _Float16* wa = &s_a[buf][srow * kUsRow];
        *(half8*)(wa + phA) = h0;
        *(half8*)(wa + phB) = h1;
        *(half8*)(wa + plA) = l0;
        *(half8*)(wa + plB) = l1;
        // the row's scale: largest magnitude of its 32 samples (this thread and
        // its partner) to [2^9, 2^10)
        float mx = 0.0f;
#pragma unroll
        for (int e = 0; e < 16; ++e) mx = fmaxf(mx, fabsf(vb[e >> 2][e & 3]));
        mx = fmaxf(mx, __shfl_xor(mx, 1, 64));
        int ex = 0;
        (void)frexpf(mx, &ex);
        const int sh = (mx > 0.0f && isfinite(mx)) ? 10 - ex : 0;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            _Float16 hi, lo;
            ub_split(ldexpf(vb[e >> 2][e & 3], sh), hi, lo);
            if (e < 8) { h0[e] = hi; l0[e] = lo; } else { h1[e - 8] = hi; l1[e - 8] = lo; }
        }
        _Float16* wb = &s_b[buf][srow * kUsRow];
        *(half8*)(wb + phA) = h0;
        *(half8*)(wb + phB) = h1;
        *(half8*)(wb + plA) = l0;
        *(half8*)(wb + plB) = l1;
        if (shalf == 0) s_sc[buf][srow] = ldexpf(1.0f, -sh - 6);
    };
    if (nst > 0) {
        fetch(k0);
        stage(0);
        __syncthreads();
    }
    for (int st = 0; st < nst; ++st) {
        const int buf = st & 1;
        if (st + 1 < nst) fetch(k0 + (int64_t)(st + 1) * kUbK);  // uniform
        if (!idle) {
            const _Float16* ra = &s_a[buf][(wr * 64 + lr) * kUsRow];
            const _Float16* rb = &s_b[buf][(wc * Cfg::WC + lr) * kUsRow];
            half8 ah[4], al[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                ah[q] = *(const half8*)(ra + 16 * q * kUsRow + rh);
                al[q] = *(const half8*)(ra + 16 * q * kUsRow + rl);
            }
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const half8 bh = *(const half8*)(rb + 16 * j * kUsRow + rh);
                const half8 bl = *(const half8*)(rb + 16 * j * kUsRow + rl);
                const float sc = s_sc[buf][wc * Cfg::WC + 16 * j + lr];
                {
                    // the four row blocks' products as independent chains, so
                    // that each matrix instruction's result is not waited on
                    // by the next one
                    V a[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        a[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bh, V{0, 0, 0, 0}, 0, 0, 0);
#pragma unroll
                    for (int i = 0; i < 4; ++i) a[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bl, a[i], 0, 0, 0);
#pragma unroll
                    for (int i = 0; i < 4; ++i) a[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[i], bh, a[i], 0, 0, 0);
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int r = 0; r < 4; ++r) tot[i][j][r] = fmaf(a[i][r], sc, tot[i][j][r]);
                }
            }
        }
        if (st + 1 < nst) stage(buf ^ 1);
        __syncthreads();
    }
    if (idle) return;
    float* __restrict__ out = part + ((size_t)kc * n_pairs + pair) * (TT * TT);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                out[(size_t)(wr * 64 + 16 * i + Mfma<float>::row(lq, r)) * TT + wc * Cfg::WC + 16 * j + lr] =
                    tot[i][j][r];
}

// Ub[l, g + 1] = fold[0] Ub[l, g + 1] + fold[1] (sum over split-K chunks of
// the partial tiles), l <= g < nd, FP64; one element per thread
template <typename T, int TT>
__global__ __launch_bounds__(256) void ubar_reduce_kernel(const T* __restrict__ part, const int2* __restrict__ pairs,
                                                          int n_pairs, int n_chunks, int nd, double* __restrict__ Ub,
                                                          int n, const double* __restrict__ fold)
{
    const int pair = blockIdx.x;
    const int2 tp = pairs[pair];
    const int e = blockIdx.y * 256 + threadIdx.x;
    if (e >= TT * TT) return;
    const int i = e / TT, j = e - i * TT;
    const int l = tp.x * TT + i, g = tp.y * TT + j;
    if (l >= nd || g >= nd || l > g) return;
    double sum = 0.0;
    for (int kc = 0; kc < n_chunks; ++kc) sum += (double)part[((size_t)kc * n_pairs + pair) * (TT * TT) + e];
    double& u = Ub[(size_t)l + (size_t)(g + 1) * n];
    u = fold[0] * u + fold[1] * sum;
}

hipError_t ubar_product(const void* Y, const void* S, int64_t rs, int64_t kmax, int nd, bool mixed,
                        const int2* pairs, int n_pairs, int n_chunks, int64_t kchunk, void* part, double* Ub, int n,
                        const double* fold, hipStream_t s, bool split16)
{
    if (kmax > rs) return hipErrorInvalidValue;
    dim3 g1(n_pairs, n_chunks), b(256);
    if (mixed) {
        if (kchunk % kUbK != 0 || rs % kUbK != 0 || kmax % kUbK != 0) return hipErrorInvalidValue;
        if (split16) {
            constexpr int TT = kUbTile16;
            hipLaunchKernelGGL(ubar_gemm_split_kernel<TT>, g1, dim3(UbSplitCfg<TT>::NT), 0, s, (const float*)Y,
                               (const float*)S, rs, nd, pairs, n_pairs, kchunk, kmax, (float*)part);
            hipLaunchKernelGGL((ubar_reduce_kernel<float, TT>), dim3(n_pairs, TT * TT / 256), b, 0, s,
                               (const float*)part, pairs, n_pairs, n_chunks, nd, Ub, n, fold);
        } else {
            constexpr int TT = kUbTile;
            hipLaunchKernelGGL(ubar_gemm_lds_kernel, g1, b, 0, s, (const float*)Y, (const float*)S, rs, nd, pairs,
                               n_pairs, kchunk, kmax, (float*)part);
            hipLaunchKernelGGL((ubar_reduce_kernel<float, TT>), dim3(n_pairs, TT * TT / 256), b, 0, s,
                               (const float*)part, pairs, n_pairs, n_chunks, nd, Ub, n, fold);
        }
    } else {
        constexpr int TT = kUbTile;
        hipLaunchKernelGGL(ubar_gemm_kernel<double>, g1, b, 0, s, (const double*)Y, (const double*)S, rs, nd,
                           pairs, n_pairs, kchunk, kmax, (double*)part);
        hipLaunchKernelGGL((ubar_reduce_kernel<double, TT>), dim3(n_pairs, TT * TT / 256), b, 0, s,
                           (const double*)part, pairs, n_pairs, n_chunks, nd, Ub, n, fold);
    }
    return hipGetLastError();
}
int ubar_tile(bool split16) { return split16 ? kUbTile16 : kUbTile; }
int ubar_wgs_per_cu(bool split16) { return split16 ? UbSplitCfg<kUbTile16>::OCC : kUbOcc; }

// diagonal of Ubar from the reverse sweep's accumulator
__global__ void set_diag_kernel(double* Ub, int n, const double* __restrict__ d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Ub[(size_t)i + (size_t)i * n] += d[i];
}

// X <- Phi(X): strict upper part zeroed, diagonal halved
__global__ void tril_half_kernel(double* X, int n)
{
    const int j = blockIdx.y;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double& v = X[(size_t)i + (size_t)j * n];
        if (i < j) v = 0.0;
        else if (i == j) v *= 0.5;
    }
}

// G[p_i, p_j] = (Cb[i, j] + Cb[j, i]) / 2 / (s_{p_i} s_{p_j})
__global__ void sigma_map_kernel(const double* __restrict__ Cb, int n, const int32_t* __restrict__ perm,
                                 const double* __restrict__ scale, double* __restrict__ G)
{
    const int j = blockIdx.y;
    const int pj = perm[j];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int pi = perm[i];
        const double c = 0.5 * (Cb[(size_t)i + (size_t)j * n] + Cb[(size_t)j + (size_t)i * n]);
        G[(size_t)pi + (size_t)pj * n] = c / (scale[pi] * scale[pj]);
    }
}

// One block per pivot position i (k = perm[i]):
//   G[k, k] -= sum_j Cs[i, j] C_p[i, j] / Sigma_kk + (ga'_i a'_i + gb'_i b'_i) / (2 Sigma_kk)
// with Cs the symmetrized C_bar and C_p[i, j] = Sigma[p_i, p_j] / (s_{p_i} s_{p_j});
// ga[k] = ga'_i / s_k, gb[k] = gb'_i / s_k.
__global__ __launch_bounds__(256) void sigma_diag_kernel(const double* __restrict__ Cb, int n,
                                                         const int32_t* __restrict__ perm,
                                                         const double* __restrict__ scale,
                                                         const double* __restrict__ Sig,
                                                         const double* __restrict__ a_std,
                                                         const double* __restrict__ b_std,
                                                         const double* __restrict__ ga_std,
                                                         const double* __restrict__ gb_std,
                                                         double* __restrict__ G, double* __restrict__ ga,
                                                         double* __restrict__ gb)
{
    __shared__ double s_r[4];
    const int i = blockIdx.x;
    const int k = perm[i];
    const int tid = threadIdx.x;
    double acc = 0.0;
    for (int j = tid; j < n; j += 256) {
        const int pj = perm[j];
        const double c = 0.5 * (Cb[(size_t)i + (size_t)j * n] + Cb[(size_t)j + (size_t)i * n]);
        const double cp = Sig[(size_t)k + (size_t)pj * n] / (scale[k] * scale[pj]);
        acc += c * cp;
    }
    acc = wave_sum_f64(acc);
    if ((tid & 63) == 0) s_r[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) {
        const double row = s_r[0] + s_r[1] + s_r[2] + s_r[3];
        const double skk = Sig[(size_t)k + (size_t)k * n];
        const double ta = isinf(a_std[i]) ? 0.0 : ga_std[i] * a_std[i];
        const double tb = isinf(b_std[i]) ? 0.0 : gb_std[i] * b_std[i];
        G[(size_t)k + (size_t)k * n] -= row / skk + (ta + tb) / (2.0 * skk);
        ga[k] = ga_std[i] / scale[k];
        gb[k] = gb_std[i] / scale[k];
    }
}

hipError_t sigma_chain(double* Ub, int n, const double* gd, double* X, hipStream_t s, int step,
                       const double* Cb, const int32_t* perm, const double* scale, const double* Sig,
                       const double* a_std, const double* b_std, const double* ga_std, const double* gb_std,
                       double* G, double* ga, double* gb)
{
    dim3 g2((n + 255) / 256, n), b(256);
    if (step == 0) {
        hipLaunchKernelGGL(set_diag_kernel, dim3((n + 255) / 256), b, 0, s, Ub, n, gd);
    } else if (step == 1) {
        hipLaunchKernelGGL(tril_half_kernel, g2, b, 0, s, X, n);
    } else {
        hipLaunchKernelGGL(sigma_map_kernel, g2, b, 0, s, Cb, n, perm, scale, G);
        hipLaunchKernelGGL(sigma_diag_kernel, dim3(n), b, 0, s, Cb, n, perm, scale, Sig, a_std, b_std, ga_std,
                           gb_std, G, ga, gb);
    }
    return hipGetLastError();
}

}  // namespace dbg
