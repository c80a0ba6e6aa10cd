// Dense integration kernel, mixed precision, with the panel contraction on the
// FP16 matrix cores (split-operand FP32 emulation).
//
// The computation is qmc_loop! (src/qmc.jl:315-444) with the point map of
// src/qmc_generators.jl:137-204 fused in, organised like the FP32 matrix-core
// kernel (qmc_dense_mfma.hip): a workgroup of four wavefronts owns 256 samples
// of one shift, the sampled dimensions are swept in panels of 128 targets,
// each panel's conditional means S = Y[:, 0:g0] U[0:g0, panel] are one dense
// contraction, and the panel is then walked in sub-blocks of 16 dimensions
// with the in-panel updates on FP32 matrix cores.
//
// What changes is the arithmetic of the panel contraction, the bulk of the
// flops.  With exact power-of-two prescales y' = 2^6 y and u' = 2^13 u, both
// operands are split into two FP16 numbers, y' = yh + yl with yh = fp16(y'),
// yl = fp16(y' - yh) (likewise u' = uh + ul), and
//     y' u' ~ yh uh + yh ul + yl uh        (yl ul, 2^-22 relative, dropped),
// three v_mfma_f32_16x16x32_f16 products that are exact in the FP32
// accumulator.  yh + yl carries y' to about 22 significant bits (FP32: 24) and
// the sum is accumulated in FP32 as in the FP32 kernel; the accumulators hold
// 2^19 S and are scaled back exactly before the walk.  This is the hi/lo FP16
// split of Ootomo and Yokota, "Recovering single precision accuracy from
// Tensor Cores while surpassing the FP32 theoretical peak performance" (IJHPCA
// 2022); instead of scaling the low parts (which would need a second set of
// accumulators) the global prescales keep the low parts of every value that
// matters in the FP16 normal range: the low part of y' is normal for
// |y| >= 2^-8 and that of u' for |u| >= 2^-15, and below those the absolute
// error stays under 2^-31 (y) and 2^-38 (u).  |u| <= 1 for the standardized
// factor and |y| <= 1023 (a mass of exp(-5e5)) keep both prescaled values in
// range; larger |y| is clamped to the FP16 range.
//
// Operand layouts (FP16, built for the 16x16x32 instruction, in which lane l
// holds row / column l % 16 and eight consecutive k of group l / 16):
//   factor  uc[j * ucs + 64 c + e]: column j (target), chunk c of 32 k,
//           e < 32 the high parts of u'[32 c + e, j], e >= 32 the low parts;
//           zero for k >= j (strict upper triangle) and for j >= n;
//   y       per workgroup slot, k-group g of 8 dimensions: 256 samples x 8
//           high parts, then 256 x 8 low parts (16 bytes per sample and part).
// A chunk of the factor (128 targets x 128 bytes) goes to LDS by direct
// global-to-LDS loads, double-buffered one chunk ahead, with the 16-byte
// granules of row r permuted by r mod 8 so that the operand reads are free of
// bank conflicts.  The y operands of the next chunk (8 KB per wavefront) go to
// LDS by direct loads as well, into the memory the walk uses for its hand-off
// buffers, so no registers hold data in flight.  The y slot is written and
// read with the non-temporal cache policy: each row passes through HBM once
// per later panel, while the factor chunks, read by every workgroup, are to
// stay in L2.
#include "dbg_device.h"

namespace dbg {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));

constexpr int kTS = 256;        // samples per workgroup (one per thread)
constexpr int kPW = 128;        // targets per panel
constexpr int kWS = 64;         // samples per wavefront
constexpr int kNSB = kWS / 16;  // 16-sample blocks per wavefront
constexpr int kNTB = kPW / 16;  // 16-target blocks per panel
constexpr int kKC = 32;         // k per factor chunk (one instruction's K)
constexpr int kSB = 16;         // walk sub-block
constexpr int kSXS = kSB + 4;   // LDS row stride of the accumulator hand-off
constexpr int kUS = kPW + 16;   // LDS row stride of the in-panel update's FP32 rows
constexpr int kChunk = kPW * 64;     // halfs per factor chunk in LDS
constexpr int kYG = 2 * kTS * 8;     // halfs per k-group of the y slot
constexpr float kYScale = 64.0f;     // 2^6
constexpr float kUScale = 8192.0f;   // 2^13
constexpr float kAccScale = 1.0f / 524288.0f;  // 2^-19
constexpr float kHalfMax = 65504.0f;
constexpr int kLoadNT = 2;           // cache policy of a direct load: non-temporal (nt)

__device__ __forceinline__ floatx4 mfma16(half8 a, half8 b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// v -> (hi, lo) with hi = fp16(v), lo = fp16(v - hi); |v| clamped to the FP16
// range (NaN kept)
__device__ __forceinline__ void split16(float v, _Float16& hi, _Float16& lo)
{
    v = fabsf(v) > kHalfMax ? copysignf(kHalfMax, v) : v;
    hi = (_Float16)v;
    lo = (_Float16)(v - (float)hi);
}

// the sub-block's 16 y values of one sample (times the prescale) -> two
// k-groups of the slot
__device__ __forceinline__ void store_y16(_Float16* __restrict__ Ys16, int gb, int t, const float (&yv)[kSB],
                                          int cnt, float scale)
{
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        half8 hi, lo;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float v = (8 * h + e < cnt) ? yv[8 * h + e] * scale : 0.0f;
            _Float16 a, b;
            split16(v, a, b);
            hi[e] = a;
            lo[e] = b;
        }
        // non-temporal: the slot is re-read from HBM by later panels, and
        // cached writes would evict the factor chunks that all workgroups share
        _Float16* dst = Ys16 + (size_t)(gb / 8 + h) * kYG + (size_t)t * 8;
        __builtin_nontemporal_store(hi, (half8*)dst);
        __builtin_nontemporal_store(lo, (half8*)(dst + kTS * 8));
    }
}

}  // namespace

int dense_split_tile() { return kTS; }
int dense_split_panel() { return kPW; }
int64_t dense_split_slot(int nd) { return (int64_t)((nd + kSB - 1) / kSB) * kSB * kTS; }  // floats

// MODE kModeIntegrate: the estimate.  kModeRecord: the same, keeping y and
// the conditional means of every sample (FP32 rows of A.rs samples in
// yrec_all / srec_all) and the per-sample weights, for the adjoint.
// kModeReverse: the adjoint's reverse sweep in flipped index order (uc, Uk,
// Ucol hold the flipped factor; the "y" of a step is sbar), as in the FP32
// matrix-core kernel, with each sample's normalised weight wt scaled to
// [1/2, 1) by a power of two so that its sbar values (all proportional to
// wt) suit the FP16 split; the scale is undone where values leave the sample
// (recorded sbar, per-index sums).  LOWER: every lower bound is -inf (the
// walk and the reverse step form only the upper bound's tail mass).
template <int GEN, int MODE, bool LOWER>
__global__ __launch_bounds__(256, 2) void dense_split_kernel(DenseArgs A, const _Float16* __restrict__ uc,
                                                            int64_t ucs, const float* __restrict__ Uk,
                                                            const float* __restrict__ Ucol,
                                                            const float* __restrict__ ak,
                                                            const float* __restrict__ bk,
                                                            const float* __restrict__ invk,
                                                            float* __restrict__ Yall,
                                                            double* __restrict__ part_m,
                                                            int* __restrict__ part_e,
                                                            float* __restrict__ yrec_all,
                                                            float* __restrict__ srec_all, DBG_POINT_PARAMS)
{
    constexpr bool REC = MODE == kModeRecord;
    constexpr bool REV = MODE == kModeReverse;
    // prescale of the split operand: y' = 2^6 y; in the reverse sweep the
    // scaled sbar are O(1) .. O(100) and go unscaled
    constexpr float kOpScale = REV ? 1.0f : kYScale;
    constexpr float kBack = REV ? 1.0f / kUScale : kAccScale;
    using V = floatx4;
    // walk: accumulator hand-off [sample][16] and y of the current sub-block
    // [k][sample]; contraction: the y chunk, per wavefront 4 k-groups x
    // (high, low) x 64 samples x 8 halfs
    __shared__ __attribute__((aligned(16))) float s_xy[kTS * kSXS + kSB * kTS];
    float* const s_x = s_xy;
    float* const s_y = s_xy + kTS * kSXS;
    _Float16* const s_yc = reinterpret_cast<_Float16*>(s_xy);
    static_assert(kTS * kKC * 2 * sizeof(_Float16) <= sizeof(s_xy), "a y chunk fits the walk buffers");
    // factor chunks of the contraction (two buffers); during the walk the same
    // memory holds the FP32 rows of the in-panel update
    __shared__ __attribute__((aligned(16))) _Float16 s_uc[2 * kChunk];
    float* const s_u = reinterpret_cast<float*>(s_uc);
    static_assert(kSB * kUS * sizeof(float) <= sizeof(s_uc), "staging fits the chunk buffers");
    // Sobol points: the sub-block's tables (low Gray-code byte per lane, high
    // part per tile), also in the chunk buffers, after the staging rows
    constexpr bool SOBT = GEN == kGenSobol && !REV;
    __shared__ double s_red[REV ? 4 * kSB * 3 : 1];
    uint32_t* const s_sob = reinterpret_cast<uint32_t*>(s_u + kSB * kUS);
    static_assert(kSB * kUS * sizeof(float) + sobol_table_words<kSB>() * sizeof(uint32_t) <= sizeof(s_uc),
                  "Sobol tables fit the chunk buffers");

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = tid >> 6;
    const int lr = lane & 15;
    const int lq = lane >> 4;
    const int n = A.n;
    const int nd = n - 1;
    const int64_t ldu = A.ldu;
    const PointArgs pts = DBG_WITH_TABLES(A.pts);

    for (int64_t it = blockIdx.x; it < A.n_tiles; it += gridDim.x) {
        const int64_t tile = A.tile0 + it;
        // records and partial sums: chunk-relative in record / reverse mode
        const int64_t rt = (REC || REV) ? it : tile;
        const int k = (int)(tile / A.tiles_per_rep);
        const int64_t jt = tile - (int64_t)k * A.tiles_per_rep;
        const int64_t j = jt * kTS + tid + 1;  // 1-based point index of this thread's sample
        const bool valid = j <= A.m;
        SobolLane sl{};
        if constexpr (SOBT) sl = sobol_lane(pts, jt * kTS + 1, j);
        _Float16* __restrict__ Ys16 = reinterpret_cast<_Float16*>(Yall + (size_t)blockIdx.x * (size_t)A.y_slot);
        float* __restrict__ yrec = (REC || REV) ? yrec_all + (size_t)rt * kTS : nullptr;
        float* __restrict__ srec = (REC || REV) ? srec_all + (size_t)rt * kTS : nullptr;

        typename Path<true, kPhiTail>::State st;
        Weight<float> w;
        float ybar = 0.0f, wt = 0.0f, unsc = 0.0f;
        // reverse: the records of the next block's first step, read ahead
        // (after each block's closing barrier; for the first block here)
        float s_nx = 0.0f, y_nx = 0.0f;
        if constexpr (REV) record_first_step(n, 0, tid, srec, yrec, A.rs, s_nx, y_nx);
        if constexpr (REV) {
            if (valid) {
                const size_t idx = (size_t)rt * kTS + tid;
                const float pmv = ((const float*)A.wm)[idx];
                if (pmv != 0.0f) {
                    const double w64 = ldexp((double)pmv, A.we[idx] - (int)A.norm[0]) * A.norm[1];
                    int e = 0;
                    wt = (float)frexp(w64, &e);
                    unsc = ldexpf(1.0f, e);  // 0 below the FP32 range, as wt itself would be
                }
            }
        } else {
            Path<true, kPhiTail>::init(st, w, A.first);
        }

        for (int g0 = 0; g0 < nd; g0 += kPW) {
            // ---- S = Y[:, 0:g0] * U[0:g0, g0+1 : g0+kPW+1] on FP16 matrix cores
            V acc[kNSB][kNTB];
#pragma unroll
            for (int sb = 0; sb < kNSB; ++sb)
#pragma unroll
                for (int tb = 0; tb < kNTB; ++tb) acc[sb][tb] = V{0, 0, 0, 0};
            const int nch = g0 / kKC;  // a multiple of 4
            if (nch > 0) {
                // lane-dependent offsets are rederived here from an opaque copy of
                // the lane index so that they are not held in registers through
                // the walk, where the register file is full
                int ln = lane;
                asm volatile("" : "+v"(ln));
                const int lr = ln & 15, lq = ln >> 4;
                // B operand reads: row 16 tb + lr of a chunk, granules lq (high)
                // and 4 + lq (low), stored at granule ^ (row mod 8)
                const int ph = (lq ^ (lr & 7)) * 8, pl = ((4 + lq) ^ (lr & 7)) * 8;
                // direct loads of a chunk: wave wv fills rows 32 wv .. 32 wv + 31,
                // eight rows (1 KB) per instruction; lane L fills granule L mod 8
                // of row L / 8 ((8 q + drow) mod 8 == drow mod 8)
                const int drow = 32 * wv + (ln >> 3);
                const int dgran = (ln & 7) ^ (drow & 7);
                // y: wave wv moves its 64 samples of k-group 4 c + q, part p
                // (1 KB) per instruction to s_yc + wv * 4096 + (2 q + p) * 512
                const _Float16* __restrict__ yd = Ys16 + (size_t)(wv * kWS + ln) * 8;
                _Float16* const ywv = s_yc + wv * 4096;
                // its operand of sample block sb: k-group lq, sample 16 sb + lr
                const _Float16* __restrict__ yr = ywv + lq * 1024 + lr * 8;
                const _Float16* __restrict__ ud = uc + (size_t)(g0 + 1 + drow) * ucs + dgran * 8;
                const _Float16* __restrict__ sr = s_uc + lr * 64;
                auto issue = [&](int c, int buf) {
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        __builtin_amdgcn_global_load_lds(ud + (size_t)(8 * q) * ucs + (size_t)c * 64,
                                                         s_uc + buf * kChunk + (32 * wv + 8 * q) * 64, 16, 0, 0);
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        // nt: the y rows stream from HBM (each is read once per
                        // panel); the non-temporal hint keeps them from evicting
                        // the factor chunks in L2
                        __builtin_amdgcn_global_load_lds(yd + (size_t)(4 * c + (q >> 1)) * kYG + (q & 1) * kTS * 8,
                                                         ywv + q * 512, 16, 0, kLoadNT);
                };
                issue(0, 0);
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                __syncthreads();
                for (int c = 0; c < nch; ++c) {
                    const int buf = c & 1;
                    half8 yh[kNSB], yl[kNSB];
#pragma unroll
                    for (int sb = 0; sb < kNSB; ++sb) {
                        yh[sb] = *(const half8*)(yr + sb * 128);
                        yl[sb] = *(const half8*)(yr + sb * 128 + 512);
                    }
                    // the next chunk's direct loads overwrite this wave's y chunk
                    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
                    if (c + 1 < nch) issue(c + 1, buf ^ 1);  // uniform
                    const _Float16* __restrict__ sb_u = sr + buf * kChunk;
#pragma unroll
                    for (int tp = 0; tp < kNTB; tp += 2) {
                        const half8 bh0 = *(const half8*)(sb_u + tp * 1024 + ph);
                        const half8 bl0 = *(const half8*)(sb_u + tp * 1024 + pl);
                        const half8 bh1 = *(const half8*)(sb_u + (tp + 1) * 1024 + ph);
                        const half8 bl1 = *(const half8*)(sb_u + (tp + 1) * 1024 + pl);
#pragma unroll
                        for (int sb = 0; sb < kNSB; ++sb) {
                            acc[sb][tp] = mfma16(yh[sb], bh0, acc[sb][tp]);
                            acc[sb][tp + 1] = mfma16(yh[sb], bh1, acc[sb][tp + 1]);
                        }
#pragma unroll
                        for (int sb = 0; sb < kNSB; ++sb) {
                            acc[sb][tp] = mfma16(yh[sb], bl0, acc[sb][tp]);
                            acc[sb][tp + 1] = mfma16(yh[sb], bl1, acc[sb][tp + 1]);
                        }
#pragma unroll
                        for (int sb = 0; sb < kNSB; ++sb) {
                            acc[sb][tp] = mfma16(yl[sb], bh0, acc[sb][tp]);
                            acc[sb][tp + 1] = mfma16(yl[sb], bh1, acc[sb][tp + 1]);
                        }
                    }
                    // the next chunk has landed (each wave its own loads; the
                    // barrier for the factor rows the other waves moved)
                    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                    __syncthreads();
                }
#pragma unroll
                for (int sb = 0; sb < kNSB; ++sb)
#pragma unroll
                    for (int tb = 0; tb < kNTB; ++tb) acc[sb][tb] *= kBack;
            }
            // ---- walk the panel in sub-blocks of 16 dimensions
            for (int b = 0; b < kNTB; ++b) {
                const int gb = g0 + kSB * b;
                if (gb >= nd) break;  // uniform
                // accumulator columns of this sub-block -> LDS [sample][col];
                // matrix-core row 4 lq + r of sample block sb is sample wv*64 + 16 sb + 4 lq + r
#pragma unroll
                for (int tb = 0; tb < kNTB; ++tb) {
                    if (tb == b) {
#pragma unroll
                        for (int sb = 0; sb < kNSB; ++sb)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                s_x[(wv * kWS + 16 * sb + 4 * lq + r) * kSXS + lr] = acc[sb][tb][r];
                    }
                }
                if constexpr (SOBT) sobol_tables<kSB, 256>(pts, gb, min(kSB, nd - gb), jt * kTS + 1, s_sob, tid);
                const bool later = b + 1 < kNTB && gb + kSB < nd;  // the panel has targets after this block
                __syncthreads();
                if (later) {
                    // factor rows gb .. gb + 15 (targets g0 + 1 .. g0 + kPW) of the
                    // in-panel update, by direct loads that the walk hides: wave
                    // wv moves rows 4 wv .. 4 wv + 3, 64 targets per instruction
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const int row = 4 * wv + (q >> 1), half = q & 1;
                        __builtin_amdgcn_global_load_lds(Uk + (size_t)(gb + row) * ldu + g0 + 1 + 64 * half + lane,
                                                         s_u + row * kUS + 64 * half, 4, 0, 0);
                    }
                }
                const int cnt = min(kSB, nd - gb);
                float yv[kSB];
                // the records (record: y from s_y, s left in s_x; reverse: sbar
                // from s_y) are stored after the wait below, which would
                // otherwise wait for them as well: they drain during the next walk
                if constexpr (REV)
                    walk_block_rev_sv<true, GEN, kTS, kSXS, kSB, false, false, LOWER>(
                        yv, pts, n, gb, cnt, k, j, tid, ybar, wt, s_x, s_y, (float*)nullptr, Ucol, A.ldc, ak, bk,
                        invk, srec, yrec, A.rs, s_red, A.a_inf != 0, unsc, s_nx, y_nx);
                else
                    walk_block_yv<true, GEN, LOWER ? kPhiTailLower : kPhiTail, kTS, kSXS, kSB, false, REC>(
                        yv, pts, gb, cnt, k, j, tid, st, w, s_x, s_y, (float*)nullptr, 0, Ucol, A.ldc, ak, bk, invk,
                        (float*)nullptr, SOBT ? s_sob : nullptr, sl);
                store_y16(Ys16, gb, tid, yv, cnt, kOpScale);
                // the direct loads above write LDS without the compiler tracking them
                if (later) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                __syncthreads();
                if constexpr (REC) {
#pragma unroll
                    for (int c = 0; c < kSB; ++c) {
                        if (c < cnt) {  // uniform
                            __builtin_nontemporal_store(s_y[c * kTS + tid], &yrec[(size_t)(gb + c) * A.rs + tid]);
                            __builtin_nontemporal_store(s_x[tid * kSXS + c], &srec[(size_t)(gb + c) * A.rs + tid]);
                        }
                    }
                }
                if constexpr (REV) {
                    // the next block's first records, in flight through the
                    // in-panel update (or the next panel's contraction)
                    if (gb + kSB < nd) record_first_step(n, gb + kSB, tid, srec, yrec, A.rs, s_nx, y_nx);
                    // sbar of flipped index gb + c replaces s in record row r - 1,
                    // r = n - 1 - (gb + c) (every record of the block was read in the walk)
#pragma unroll
                    for (int c = 0; c < kSB; ++c) {
                        if (c < cnt)  // uniform
                            __builtin_nontemporal_store(s_y[c * kTS + tid] * unsc,
                                                        &srec[(size_t)(n - 2 - (gb + c)) * A.rs + tid]);
                    }
                    // per-index sums of the four waves -> the global accumulators
                    if (tid < 3 * cnt) {
                        const int cc = tid / 3, q = tid - 3 * cc;
                        double v = 0.0;
#pragma unroll
                        for (int ww = 0; ww < kTS / 64; ++ww) v += s_red[(ww * kSB + cc) * 3 + q];
                        double* dst = q == 0 ? A.ga : (q == 1 ? A.gb : A.gd);
                        atomicAdd(dst + (n - 1 - (gb + cc)), v);
                    }
                }
                // later targets of the panel += Y_sub * U[sub, later] (FP32 matrix cores)
                if (later) {
#pragma unroll
                    for (int ks = 0; ks < kSB / 4; ++ks) {
                        float av[kNSB];
#pragma unroll
                        for (int sb = 0; sb < kNSB; ++sb) av[sb] = s_y[(4 * ks + lq) * kTS + wv * kWS + 16 * sb + lr];
                        const float* __restrict__ ub = s_u + (4 * ks + lq) * kUS + lr;
#pragma unroll
                        for (int tb = 0; tb < kNTB; ++tb) {
                            if (tb > b) {
                                const float bv = ub[16 * tb];
#pragma unroll
                                for (int sb = 0; sb < kNSB; ++sb)
                                    acc[sb][tb] = Mfma<float>::mac(av[sb], bv, acc[sb][tb]);
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }
        if constexpr (REV) {
            // index 0: s_0 = 0, no sbar to pass on
            RevOut o{0.0, 0.0, 0.0};
            (void)rev_step<true, GEN, LOWER>(pts, n, 0, k, j, ybar, wt, 0.0f, n >= 2 ? yrec[tid] : 0.0f, ak[0],
                                             bk[0], invk[0], o);
            const double a0 = wave_sum_f64(o.ga * unsc), b0 = wave_sum_f64(o.gb * unsc),
                         d0 = wave_sum_f64(o.gd * unsc);
            if (lane == 0) {
                atomicAdd(A.ga, a0);
                atomicAdd(A.gb, b0);
                atomicAdd(A.gd, d0);
            }
            continue;
        }
        const double pm = valid ? (double)w.pm : 0.0;
        if constexpr (REC) {
            ((float*)A.wm)[(size_t)rt * kTS + tid] = valid ? w.pm : 0.0f;
            A.we[(size_t)rt * kTS + tid] = w.pe;
        }
        double S;
        int E;
        block_scaled_sum<256>(pm, w.pe, S, E);
        if (tid == 0) {
            part_m[rt] = S;
            part_e[rt] = E;
        }
    }
}

// uc[j * ucs + 64 c + e] from the FP32 column-major factor (ldc = n): high
// and low FP16 parts of 2^13 U[32 c + e, j] for 32 c + e < j < n, else zero.
// One block per target column j < ncol.
__global__ void pack_split16_kernel(const float* __restrict__ Ucol, int n, _Float16* __restrict__ uc, int64_t ucs)
{
    const int jcol = blockIdx.x;
    for (int64_t idx = threadIdx.x; idx < ucs / 2; idx += blockDim.x) {
        const int c = (int)(idx / 32), e = (int)(idx % 32);
        const int l = 32 * c + e;
        const float v = (jcol < n && l < jcol) ? Ucol[(size_t)jcol * n + l] * kUScale : 0.0f;
        _Float16 hi, lo;
        split16(v, hi, lo);
        uc[(size_t)jcol * ucs + 64 * c + e] = hi;
        uc[(size_t)jcol * ucs + 64 * c + 32 + e] = lo;
    }
}

int64_t split16_col_stride(int n) { return (int64_t)((n + kKC - 1) / kKC) * 64; }
int split16_cols(int n) { return n + kPW + 32; }

hipError_t pack_split16(const float* Ucol, int n, void* uc, hipStream_t s)
{
    const int64_t ucs = split16_col_stride(n);
    hipLaunchKernelGGL(pack_split16_kernel, dim3(split16_cols(n)), dim3(256), 0, s, Ucol, n, (_Float16*)uc, ucs);
    return hipGetLastError();
}

template <int GEN, int MODE, bool LOWER = false>
static void launch_s(const DenseArgs& A, const void* uc, int64_t ucs, int grid, hipStream_t s)
{
    hipLaunchKernelGGL((dense_split_kernel<GEN, MODE, LOWER>), dim3(grid), dim3(256), 0, s, A, (const _Float16*)uc, ucs,
                       (const float*)A.Uk, (const float*)A.Ucol, (const float*)A.ak, (const float*)A.bk,
                       (const float*)A.invk, (float*)A.Y, A.part_m, A.part_e, (float*)A.yrec, (float*)A.srec,
                       DBG_POINT_ARGS(A.pts));
}

hipError_t launch_dense_split(const DenseArgs& A, int gen, const void* uc, int grid, hipStream_t s)
{
    const int64_t ucs = split16_col_stride(A.n);
    const bool lower = A.a_inf != 0;
    if (gen == kGenRichtmyer) {
        if (A.mode == kModeRecord) {
            if (lower) launch_s<kGenRichtmyer, kModeRecord, true>(A, uc, ucs, grid, s);
            else launch_s<kGenRichtmyer, kModeRecord>(A, uc, ucs, grid, s);
        } else if (A.mode == kModeReverse) {
            if (lower) launch_s<kGenRichtmyer, kModeReverse, true>(A, uc, ucs, grid, s);
            else launch_s<kGenRichtmyer, kModeReverse>(A, uc, ucs, grid, s);
        } else {
            if (lower) launch_s<kGenRichtmyer, kModeIntegrate, true>(A, uc, ucs, grid, s);
            else launch_s<kGenRichtmyer, kModeIntegrate>(A, uc, ucs, grid, s);
        }
    } else {
        if (A.mode == kModeRecord) {
            if (lower) launch_s<kGenSobol, kModeRecord, true>(A, uc, ucs, grid, s);
            else launch_s<kGenSobol, kModeRecord>(A, uc, ucs, grid, s);
        } else if (A.mode == kModeReverse) {
            if (lower) launch_s<kGenSobol, kModeReverse, true>(A, uc, ucs, grid, s);
            else launch_s<kGenSobol, kModeReverse>(A, uc, ucs, grid, s);
        } else {
            if (lower) launch_s<kGenSobol, kModeIntegrate, true>(A, uc, ucs, grid, s);
            else launch_s<kGenSobol, kModeIntegrate>(A, uc, ucs, grid, s);
        }
    }
    return hipGetLastError();
}

}  // namespace dbg
