// Dense integration kernel, matrix-core form (large d).
//
// The same computation as qmc_loop! (src/qmc.jl:315-444) with the point map
// of src/qmc_generators.jl:137-204 fused in, organised around the MI355X
// matrix cores:
//
//   * a workgroup (4 wavefronts) owns TS samples of one shift; its sampled
//     y values live in a per-block workspace slot, layout [dimension][sample];
//   * the sampled dimensions are swept in panels of PW targets.  For each
//     panel the conditional means S (TS x PW) of all earlier dimensions are one
//     dense contraction S = Y[:, 0:g0] * U[0:g0, panel] on v_mfma_*_16x16x4,
//     with the accumulators held in registers (the outer mul! of :354);
//   * the panel is then walked in sub-blocks of 16 dimensions: the
//     sub-block's 16 accumulator columns go through LDS to a lane-per-sample
//     layout, each lane walks its sample through the 16 dimensions in
//     registers (point, Phi^{-1}, the in-block triangle of :390-403, Phi,
//     weight), and the 16 new y columns update the panel's later targets with
//     one more K = 16 matrix-core contraction (the incremental mul! of :374).
//
// Operands of the contraction are read straight from global memory: Y rows
// are private to the workgroup (coalesced 64-byte segments), U rows are shared
// by all workgroups and stay in L2 / MALL.
#include <type_traits>

#include "dbg_device.h"

namespace dbg {

constexpr int kSB = 16;      // walk sub-block (dimensions)
constexpr int kSXS = kSB + 4;  // LDS row stride of the accumulator hand-off

template <bool MIXED, bool NARROW = false>
struct MfmaCfg;
// TS samples x PW targets per workgroup; PFY / PFU: depths of the operand
// rings of the panel contraction (Y rows stream from HBM, U rows are shared by
// every workgroup and come from L2); OCC: waves per SIMD the register budget
// is cut for.
// ULDS: U rows of the panel contraction go through LDS in chunks of kUC rows,
// loaded once per workgroup (instead of once per wavefront) one chunk ahead.
template <>
struct MfmaCfg<true, false> {   // FP32: 256 samples x 128 targets, 128 accumulators per lane
    static constexpr int TS = 256, PW = 128, PFY = 4, PFU = 2, OCC = 2;
    static constexpr bool ULDS = true;
};
template <>
struct MfmaCfg<false, false> {  // FP64: 256 samples x 64 targets, 128 accumulator registers per lane
    // (one sample per thread: every wavefront walks; the narrower panel costs
    // twice the y re-reads of 128 targets, which the walk-bound kernel hides)
    static constexpr int TS = 256, PW = 64, PFY = 4, PFU = 4, OCC = 1;
    static constexpr bool ULDS = false;
};
template <>
struct MfmaCfg<false, true> {   // FP64, 128 samples x 128 targets: the re-evaluation of a few
    // tiles, bound by one tile's latency (half the panels of the form above)
    static constexpr int TS = 128, PW = 128, PFY = 4, PFU = 4, OCC = 1;
    static constexpr bool ULDS = false;
};
constexpr int kUC = 16;  // U rows per LDS chunk (four K-steps)

// Sample order inside a wavefront's WS samples: the matrix-core row lr of
// sample block sb is sample NSB * lr + sb, so the NSB operands a lane needs
// from one Y row are adjacent (16 bytes: one vector load).
template <typename T, int NSB>
struct alignas(16) YVec {
    T v[NSB];
};

// Y rows k0 + lq of the lane's NSB samples; non-temporal loads (the rows
// stream from HBM once per panel and would otherwise crowd the factor rows
// out of L2)
template <typename T, int NSB>
__device__ __forceinline__ void load_y(T (&av)[NSB], const T* __restrict__ ya, int64_t ys, int k0)
{
    static_assert(sizeof(YVec<T, NSB>) % 16 == 0, "16-byte loads per lane and K-step");
    typedef T VT __attribute__((ext_vector_type(NSB)));
    const VT v = __builtin_nontemporal_load((const VT*)(ya + (size_t)k0 * ys));
#pragma unroll
    for (int sb = 0; sb < NSB; ++sb) av[sb] = v[sb];
}

// the same operands from the LDS copy of a sub-block's y
template <typename T, int NSB>
__device__ __forceinline__ void load_y_lds(T (&av)[NSB], const T* __restrict__ ya, int64_t ys, int k0)
{
    const YVec<T, NSB> v = *(const YVec<T, NSB>*)(ya + (size_t)k0 * ys);
#pragma unroll
    for (int sb = 0; sb < NSB; ++sb) av[sb] = v.v[sb];
}

// U rows k0 + lq of the panel's NTB target blocks
template <typename T, int NTB>
__device__ __forceinline__ void load_u(T (&bv)[NTB], const T* __restrict__ ub, int64_t ldu, int k0)
{
#pragma unroll
    for (int tb = 0; tb < NTB; ++tb) bv[tb] = ub[(size_t)k0 * ldu + 16 * tb];
}

int dense_mfma_tile(bool mixed) { return mixed ? MfmaCfg<true>::TS : MfmaCfg<false>::TS; }
int dense_mfma_panel(bool mixed) { return mixed ? MfmaCfg<true>::PW : MfmaCfg<false>::PW; }
int dense_mfma_refine_tile() { return MfmaCfg<false, true>::TS; }
int dense_mfma_refine_panel() { return MfmaCfg<false, true>::PW; }

// MODE kModeIntegrate: the estimate.  kModeRecord: the same, keeping y and
// the conditional means of every sample (rows of A.rs) and the per-sample
// weights, for the adjoint.  kModeReverse: the reverse sweep in flipped index
// order (Uk / Ucol hold the flipped factor, the "y" of a step is sbar); see
// walk_block_rev.
template <bool MIXED, int GEN, int FORM, int MODE, bool NARROW = false,
          typename T = typename std::conditional<MIXED, float, double>::type>
__global__ __launch_bounds__(256, (MfmaCfg<MIXED, NARROW>::OCC)) void dense_mfma_kernel(DenseArgs A, const T* __restrict__ Uk,
                                                         const T* __restrict__ Ucol,
                                                         const T* __restrict__ ak,
                                                         const T* __restrict__ bk,
                                                         const T* __restrict__ invk,
                                                         T* __restrict__ Yall,
                                                         double* __restrict__ part_m,
                                                         int* __restrict__ part_e,
                                                         T* __restrict__ srec_all,
                                                         const T* __restrict__ yrec_all,
                                                         DBG_POINT_PARAMS)
{
    using V = typename Mfma<T>::V;
    constexpr int TS = MfmaCfg<MIXED, NARROW>::TS;
    constexpr int PW = MfmaCfg<MIXED, NARROW>::PW;
    // the Sobol walk has no room in the FP32 register budget for the deeper Y ring
    constexpr int kPFY = (MIXED && GEN == kGenSobol) ? 2 : MfmaCfg<MIXED, NARROW>::PFY;
    constexpr int kPFU = MfmaCfg<MIXED, NARROW>::PFU;
    constexpr int kUnroll = kPFY > kPFU ? kPFY : kPFU;
    static_assert(kUnroll % kPFY == 0 && kUnroll % kPFU == 0, "ring depths must divide the unroll");
    constexpr int WS = TS / 4;         // samples per wavefront
    constexpr int NSB = WS / 16;       // 16-sample blocks per wavefront
    constexpr int NTB = PW / 16;       // 16-target blocks per panel
    constexpr bool REV = MODE == kModeReverse;
    constexpr bool REC = MODE == kModeRecord;

    __shared__ T s_x[TS * kSXS];       // accumulator hand-off, [sample][16]
    __shared__ T s_y[kSB * TS];        // y of the current sub-block, [k][sample]
    __shared__ double s_red[REV ? 4 * kSB * 3 : 1];
    constexpr bool ULDS = MfmaCfg<MIXED, NARROW>::ULDS;
    // the in-panel update's factor rows staged in LDS during the walk (not for
    // Sobol: its LDS tables leave no registers for the staging addresses)
    constexpr bool UPDL = ULDS && GEN != kGenSobol;
    constexpr int kUS = PW + 16;       // LDS row stride of a U chunk (conflict-free reads)
    __shared__ T s_u[ULDS ? 2 * kUC * kUS : 1];
    constexpr bool SOBT = GEN == kGenSobol && !REV;  // walk points from LDS Sobol tables
    __shared__ uint32_t s_sob[SOBT ? sobol_table_words<kSB>() : 1];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = tid >> 6;
    const int lr = lane & 15;          // row/column inside a 16x16 block
    const int lq = lane >> 4;          // k (operands) or row quad (result)
    const int n = A.n;
    const int nd = n - 1;
    const int64_t ldu = A.ldu;
    const int64_t ys = REC ? A.rs : TS;  // row stride of the y workspace
    const bool walker = tid < TS;
    const PointArgs pts = DBG_WITH_TABLES(A.pts);

    for (int64_t it = blockIdx.x; it < A.n_tiles; it += gridDim.x) {
        const int64_t tile = A.tile_list ? A.tile_list[it] : A.tile0 + it;
        if (tile < 0) continue;  // uniform: one entry per workgroup
        // records and partial sums: chunk-relative in record / reverse mode
        const int64_t rt = (REC || REV) ? it : tile;
        const int k = (int)(tile / A.tiles_per_rep);
        const int64_t jt = tile - (int64_t)k * A.tiles_per_rep;
        const int64_t j = jt * TS + tid + 1;  // 1-based point index of the walker's sample
        const bool valid = walker && j <= A.m;
        SobolLane sl{};
        if constexpr (SOBT) sl = sobol_lane(pts, jt * TS + 1, j);
        T* __restrict__ Ys = REC ? Yall + (size_t)rt * TS : Yall + (size_t)blockIdx.x * (size_t)A.y_slot;
        T* __restrict__ srec = (REC || REV) ? srec_all + (size_t)rt * TS : nullptr;
        const T* __restrict__ yrec = REV ? yrec_all + (size_t)rt * TS : nullptr;

        typename Path<MIXED, FORM>::State st;
        Weight<T> w;
        T ybar = T(0), wt = T(0);
        if constexpr (REV) {
            if (valid) {
                const size_t idx = (size_t)rt * TS + tid;
                const T pmv = ((const T*)A.wm)[idx];
                wt = pmv != T(0) ? (T)(ldexp((double)pmv, A.we[idx] - (int)A.norm[0]) * A.norm[1]) : T(0);
            }
        } else {
            Path<MIXED, FORM>::init(st, w, A.first);
        }

        for (int g0 = 0; g0 < nd; g0 += PW) {
            // ---- S = Y[:, 0:g0] * U[0:g0, g0+1 : g0+PW+1]
            V acc[NSB][NTB];
#pragma unroll
            for (int sb = 0; sb < NSB; ++sb)
#pragma unroll
                for (int tb = 0; tb < NTB; ++tb) acc[sb][tb] = V{0, 0, 0, 0};
            if constexpr (ULDS) {
                // Y rows: a ring of kPFY K-steps in registers (private to the
                // wavefront).  U rows: chunks of kUC rows x PW targets in LDS,
                // double-buffered; each thread stages 8 targets of one row of
                // the next chunk in registers while the current chunk runs.
                static_assert(PW == 128 && kUC == 16, "staging: 256 threads x 8 = kUC x PW");
                static_assert(4 % kPFY == 0, "Y ring turns once per chunk");
                const T* __restrict__ ya = Ys + (size_t)lq * ys + wv * WS + NSB * lr;
                const int srow = tid >> 4, scol = (tid & 15) * 8;
                const T* __restrict__ us = Uk + (size_t)srow * ldu + g0 + 1 + scol;
                T* __restrict__ sw = s_u + srow * kUS + scol;
                const T* __restrict__ sr = s_u + lq * kUS + lr;
                T av[kPFY][NSB], stg[8];
                const int nch = g0 / kUC;
                if (nch > 0) {
#pragma unroll
                    for (int q = 0; q < kPFY - 1; ++q) load_y<T, NSB>(av[q], ya, ys, 4 * q);
#pragma unroll
                    for (int e = 0; e < 8; ++e) stg[e] = us[e];
#pragma unroll
                    for (int e = 0; e < 8; ++e) sw[e] = stg[e];
                    __syncthreads();
                }
                for (int c = 0; c < nch; ++c) {
                    const T* __restrict__ sb_u = sr + (c & 1) * (kUC * kUS);
                    if (c + 1 < nch) {  // uniform
#pragma unroll
                        for (int e = 0; e < 8; ++e) stg[e] = us[(size_t)(c + 1) * kUC * ldu + e];
                    }
#pragma unroll
                    for (int ks = 0; ks < 4; ++ks) {
                        const int k0 = (4 * c + ks) * 4;
                        load_y<T, NSB>(av[(ks + kPFY - 1) % kPFY], ya, ys, min(k0 + 4 * (kPFY - 1), g0 - 4));
                        T bv[NTB];
#pragma unroll
                        for (int tb = 0; tb < NTB; ++tb) bv[tb] = sb_u[4 * ks * kUS + 16 * tb];
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int sb = 0; sb < NSB; ++sb)
#pragma unroll
                            for (int tb = 0; tb < NTB; ++tb)
                                acc[sb][tb] = Mfma<T>::mac(av[ks % kPFY][sb], bv[tb], acc[sb][tb]);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                    if (c + 1 < nch) {
                        T* __restrict__ swb = sw + ((c + 1) & 1) * (kUC * kUS);
#pragma unroll
                        for (int e = 0; e < 8; ++e) swb[e] = stg[e];
                    }
                    __syncthreads();
                }
            } else {
                // operands of K-step k0 (4 dimensions) are loaded kPFY - 1 (Y)
                // and kPFU - 1 (U) steps ahead into rings of register buffers,
                // so the loads of later steps are in flight while the matrix
                // cores run
                const T* __restrict__ ya = Ys + (size_t)lq * ys + wv * WS + NSB * lr;
                const T* __restrict__ ub = Uk + (size_t)lq * ldu + g0 + 1 + lr;
                // g0 is a multiple of PW, hence of 4 kUnroll: no partial ring
                // turn.  Prefetches past the panel start re-read the last
                // K-step (unconditional loads keep the wait counts exact).
                static_assert(PW % (4 * kUnroll) == 0, "panel width must be a multiple of the ring");
                T av[kPFY][NSB], bv[kPFU][NTB];
                if (g0 > 0) {
#pragma unroll
                    for (int q = 0; q < kPFY - 1; ++q) load_y<T, NSB>(av[q], ya, ys, 4 * q);
#pragma unroll
                    for (int q = 0; q < kPFU - 1; ++q) load_u<T, NTB>(bv[q], ub, ldu, 4 * q);
                }
                for (int k0 = 0; k0 < g0; k0 += 4 * kUnroll) {
#pragma unroll
                    for (int q = 0; q < kUnroll; ++q) {
                        load_y<T, NSB>(av[(q + kPFY - 1) % kPFY], ya, ys, min(k0 + 4 * (q + kPFY - 1), g0 - 4));
                        load_u<T, NTB>(bv[(q + kPFU - 1) % kPFU], ub, ldu, min(k0 + 4 * (q + kPFU - 1), g0 - 4));
                        // keep the loads ahead of this step's matrix-core work:
                        // left to itself the scheduler interleaves them and
                        // consumes loads a few instructions after issue
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int sb = 0; sb < NSB; ++sb)
#pragma unroll
                            for (int tb = 0; tb < NTB; ++tb)
                                acc[sb][tb] = Mfma<T>::mac(av[q % kPFY][sb], bv[q % kPFU][tb], acc[sb][tb]);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
            }
            // ---- walk the panel in sub-blocks of 16 dimensions
            for (int b = 0; b < NTB; ++b) {
                const int gb = g0 + kSB * b;
                if (gb >= nd) break;  // uniform
                // accumulator columns of this sub-block -> LDS [sample][col]
#pragma unroll
                for (int tb = 0; tb < NTB; ++tb) {
                    if (tb == b) {
#pragma unroll
                        for (int sb = 0; sb < NSB; ++sb)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                s_x[(wv * WS + NSB * Mfma<T>::row(lq, r) + sb) * kSXS + lr] = acc[sb][tb][r];
                    }
                }
                if constexpr (SOBT) sobol_tables<kSB, 256>(pts, gb, min(kSB, nd - gb), jt * TS + 1, s_sob, tid);
                const bool later = b + 1 < NTB && gb + kSB < nd;  // the panel has targets after this block
                __syncthreads();
                if constexpr (UPDL) {
                    // factor rows gb .. gb + 15 (targets g0 + 1 .. g0 + PW) of the update after
                    // the walk go to LDS (the contraction's chunk buffer, idle until then) by
                    // direct loads issued before the walk, so the walk hides their latency
                    // (the barrier after the walk waits for them): wave wv moves rows
                    // 4 wv .. 4 wv + 3, 64 targets per instruction
                    if (later) {
#pragma unroll
                        for (int q = 0; q < 8; ++q) {
                            const int row = 4 * wv + (q >> 1), half = q & 1;
                            __builtin_amdgcn_global_load_lds(Uk + (size_t)(gb + row) * ldu + g0 + 1 + 64 * half + lane,
                                                             s_u + row * kUS + 64 * half, 4, 0, 0);
                        }
                    }
                }
                const int cnt = min(kSB, nd - gb);
                if constexpr (REV) {
                    if (walker)
                        walk_block_rev<MIXED, GEN, TS, kSXS, kSB>(pts, n, gb, cnt, k, j, tid, ybar, wt, s_x, s_y, Ys,
                                                                  Ucol, A.ldc, ak, bk, invk, srec, yrec, A.rs,
                                                                  s_red, A.a_inf != 0);
                } else {
                    if (walker)
                        walk_block<MIXED, GEN, FORM, TS, kSXS, kSB>(pts, gb, cnt, k, j, tid, st, w, s_x, s_y, Ys, ys,
                                                                     Ucol, A.ldc, ak, bk, invk, srec,
                                                                     SOBT ? s_sob : nullptr, sl);
                }
                if constexpr (UPDL) {
                    // the direct loads above write LDS without the compiler tracking them
                    if (later) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                }
                __syncthreads();
                if constexpr (REV) {
                    // per-index sums of the four waves -> the global accumulators
                    if (tid < 3 * cnt) {
                        const int cc = tid / 3, q = tid - 3 * cc;
                        double v = 0.0;
#pragma unroll
                        for (int ww = 0; ww < TS / 64; ++ww) v += s_red[(ww * kSB + cc) * 3 + q];  // walker waves
                        double* dst = q == 0 ? A.ga : (q == 1 ? A.gb : A.gd);
                        atomicAdd(dst + (n - 1 - (gb + cc)), v);
                    }
                }
                // later targets of the panel += Y_sub * U[sub, later]
                if (later) {
#pragma unroll
                    for (int ks = 0; ks < kSB / 4; ++ks) {
                        T av[NSB];
                        load_y_lds<T, NSB>(av, s_y + (size_t)lq * TS + wv * WS + NSB * lr, TS, 4 * ks);
                        const T* __restrict__ ub = UPDL ? s_u + (4 * ks + lq) * kUS + lr
                                                        : Uk + (size_t)(gb + 4 * ks + lq) * ldu + g0 + 1 + lr;
#pragma unroll
                        for (int tb = 0; tb < NTB; ++tb) {
                            if (tb > b) {
                                const T bv = ub[16 * tb];
#pragma unroll
                                for (int sb = 0; sb < NSB; ++sb) acc[sb][tb] = Mfma<T>::mac(av[sb], bv, acc[sb][tb]);
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
            if (walker)
                (void)rev_step<MIXED, GEN>(pts, n, 0, k, j, ybar, wt, T(0), n >= 2 ? yrec[tid] : T(0), ak[0], bk[0],
                                           invk[0], o);
            const double a0 = wave_sum_f64(o.ga), b0 = wave_sum_f64(o.gb), d0 = wave_sum_f64(o.gd);
            if (lane == 0 && walker) {
                atomicAdd(A.ga, a0);
                atomicAdd(A.gb, b0);
                atomicAdd(A.gd, d0);
            }
        } else {
            const double pm = valid ? (double)w.pm : 0.0;
            if constexpr (REC) {
                if (walker) {
                    ((T*)A.wm)[(size_t)rt * TS + tid] = valid ? w.pm : T(0);
                    A.we[(size_t)rt * TS + tid] = w.pe;
                }
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
}

template <bool MIXED, int GEN, int FORM, int MODE = kModeIntegrate, bool NARROW = false>
static void launch_m(const DenseArgs& A, int grid, hipStream_t s)
{
    using T = typename std::conditional<MIXED, float, double>::type;
    T* Y = (MODE == kModeRecord) ? (T*)A.yrec : (T*)A.Y;
    hipLaunchKernelGGL((dense_mfma_kernel<MIXED, GEN, FORM, MODE, NARROW>), dim3(grid), dim3(256), 0, s, A,
                       (const T*)A.Uk, (const T*)A.Ucol, (const T*)A.ak, (const T*)A.bk,
                       (const T*)A.invk, Y, A.part_m, A.part_e, (T*)A.srec, (const T*)A.yrec,
                       DBG_POINT_ARGS(A.pts));
}

hipError_t launch_dense_mfma(const DenseArgs& A, bool mixed, int gen, int form, int grid,
                             hipStream_t s)
{
    if (mixed) {
        // a lower orthant takes the walk that forms only the upper bound's tail mass
        if (gen == kGenRichtmyer) {
            if (A.a_inf) launch_m<true, kGenRichtmyer, kPhiTailLower>(A, grid, s);
            else launch_m<true, kGenRichtmyer, kPhiTail>(A, grid, s);
        } else {
            if (A.a_inf) launch_m<true, kGenSobol, kPhiTailLower>(A, grid, s);
            else launch_m<true, kGenSobol, kPhiTail>(A, grid, s);
        }
    } else if (gen == kGenRichtmyer) {
        if (form == kPhiErf)
            launch_m<false, kGenRichtmyer, kPhiErf>(A, grid, s);
        else if (form == kPhiErfc)
            launch_m<false, kGenRichtmyer, kPhiErfc>(A, grid, s);
        else
            launch_m<false, kGenRichtmyer, kPhiTail>(A, grid, s);
    } else {
        if (form == kPhiErf)
            launch_m<false, kGenSobol, kPhiErf>(A, grid, s);
        else if (form == kPhiErfc)
            launch_m<false, kGenSobol, kPhiErfc>(A, grid, s);
        else
            launch_m<false, kGenSobol, kPhiTail>(A, grid, s);
    }
    return hipGetLastError();
}

hipError_t launch_dense_mfma_refine(const DenseArgs& A, int gen, int grid, hipStream_t s)
{
    if (gen == kGenRichtmyer)
        launch_m<false, kGenRichtmyer, kPhiTail, kModeIntegrate, true>(A, grid, s);
    else
        launch_m<false, kGenSobol, kPhiTail, kModeIntegrate, true>(A, grid, s);
    return hipGetLastError();
}

hipError_t launch_dense_adjoint(const DenseArgs& A, bool mixed, int gen, int grid, hipStream_t s)
{
    const bool rev = A.mode == kModeReverse;
    if (mixed) {
        if (gen == kGenRichtmyer) {
            if (rev) launch_m<true, kGenRichtmyer, kPhiTail, kModeReverse>(A, grid, s);
            else launch_m<true, kGenRichtmyer, kPhiTail, kModeRecord>(A, grid, s);
        } else {
            if (rev) launch_m<true, kGenSobol, kPhiTail, kModeReverse>(A, grid, s);
            else launch_m<true, kGenSobol, kPhiTail, kModeRecord>(A, grid, s);
        }
    } else {
        if (gen == kGenRichtmyer) {
            if (rev) launch_m<false, kGenRichtmyer, kPhiTail, kModeReverse>(A, grid, s);
            else launch_m<false, kGenRichtmyer, kPhiTail, kModeRecord>(A, grid, s);
        } else {
            if (rev) launch_m<false, kGenSobol, kPhiTail, kModeReverse>(A, grid, s);
            else launch_m<false, kGenSobol, kPhiTail, kModeRecord>(A, grid, s);
        }
    }
    return hipGetLastError();
}

}  // namespace dbg
