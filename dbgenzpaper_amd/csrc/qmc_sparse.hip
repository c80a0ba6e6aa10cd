// Sparse / banded integration kernel: qmc_loop!(::QMCDataSparse)
// (src/qmc.jl:618-718) and SparseQMC.qmc_loop! (src/qmc_sparse.jl:159-262).
//
// One sample per lane, 256 points of one shift per block.  For each sampled
// dimension the conditional mean of the next dimension is the dot of the
// stored column of triu(U, 1) with the sample's earlier y values, in CSC row
// order.  Row indices and values are the same for every lane (scalar loads).
//
// Where y values live: a banded factor only ever reads the last few rows, so
// when the pattern's reach (largest distance between the newest dimension and
// a row it reads) fits, the rows sit in an LDS ring of RW rows x 256 lanes,
// row r in slot r mod RW, overwritten RW dimensions later, after its last
// reader.  Each lane reads back only its own column, so no barrier is needed,
// and the block needs no global workspace.  A wider pattern keeps y in the
// block's global workspace slot (RW = 0).  No contraction here: with the ring
// the kernel is bound by the special functions.
#include <type_traits>

#include "dbg_device.h"

namespace dbg {

constexpr int kTileS = 256;
int sparse_tile() { return kTileS; }

// LDS ring sizes: RW x 256 x sizeof(T) stays within 64 KB per block
constexpr int kRingMax32 = 64;
constexpr int kRingMax64 = 32;

int sparse_ring(bool mixed, int reach)
{
    const int cap = mixed ? kRingMax32 : kRingMax64;
    for (int rw = 16; rw <= cap; rw *= 2)
        if (reach < rw) return rw;
    return 0;
}

template <bool MIXED, int GEN, int FORM, int RW>
__global__ __launch_bounds__(kTileS) void sparse_kernel(SparseArgs A, DBG_POINT_PARAMS)
{
    const PointArgs pts = DBG_WITH_TABLES(A.pts);
    using T = typename std::conditional<MIXED, float, double>::type;
    using PT = Path<MIXED, FORM>;
    static_assert(RW == 0 || (RW & (RW - 1)) == 0, "ring size is a power of two");
    __shared__ T ring[RW > 0 ? RW : 1][kTileS];
    const int tid = threadIdx.x;
    const int nd = A.n - 1;
    const int64_t* __restrict__ colptr = A.colptr;
    const int32_t* __restrict__ rowval = A.rowval;
    const T* __restrict__ nzval = (const T*)A.nzval;
    const T* __restrict__ ak = (const T*)A.ak;
    const T* __restrict__ bk = (const T*)A.bk;
    const T* __restrict__ invk = (const T*)A.invk;
    T* __restrict__ Ys = (RW > 0) ? nullptr : (T*)A.Y + (size_t)blockIdx.x * (size_t)A.y_slot;

    for (int64_t tile = blockIdx.x; tile < A.n_tiles; tile += gridDim.x) {
        const int k = (int)(tile / A.tiles_per_rep);
        const int64_t jt = tile - (int64_t)k * A.tiles_per_rep;
        const int64_t j = jt * kTileS + tid + 1;
        const bool valid = j <= A.m;
        typename PT::State st;
        Weight<T> w;
        PT::init(st, w, A.first);
        for (int g = 0; g < nd; ++g) {
            // the SparseQMC module uses the shifted point without the tent map
            T y;
            if constexpr (MIXED) {
                const uint32_t u = point_u32<GEN>(pts, g, k, j);
                float wf, wc;
                if (A.tent) tent_w(u, wf, wc);
                else plain_w(u, wf, wc);
                y = tail_sample_f<true>(st, wf, wc);
            } else {
                const double u = point_u<GEN>(pts, g, k, j);
                y = A.tent ? path_sample<MIXED, FORM>(st, u) : PT::sample_w(st, u);
            }
            if constexpr (RW > 0) ring[g & (RW - 1)][tid] = y;
            else Ys[(size_t)g * kTileS + tid] = y;
            T sv = T(0);
            const int64_t p1 = colptr[g + 2];
            for (int64_t p = colptr[g + 1]; p < p1; ++p) {
                const int r = rowval[p];
                T yr;
                if constexpr (RW > 0) yr = ring[r & (RW - 1)][tid];
                else yr = Ys[(size_t)r * kTileS + tid];
                sv = fma(yr, nzval[p], sv);
            }
            if constexpr (MIXED)  // Path<true>::phi with the special functions as calls
                tail_from_z<true>(st, (ak[g + 1] - sv) * invk[g + 1], (bk[g + 1] - sv) * invk[g + 1]);
            else
                PT::phi(st, sv, ak[g + 1], bk[g + 1], invk[g + 1]);
            PT::weigh(w, st);
        }
        const double pm = valid ? (double)w.pm : 0.0;
        double S;
        int E;
        block_scaled_sum<kTileS>(pm, w.pe, S, E);
        if (tid == 0) {
            A.part_m[tile] = S;
            A.part_e[tile] = E;
        }
    }
}

using SparseKernel = void (*)(SparseArgs, DBG_POINT_PARAMS);

template <bool MIXED, int GEN, int FORM>
SparseKernel pick_ring(int ring)
{
    switch (ring) {
    case 0: return sparse_kernel<MIXED, GEN, FORM, 0>;
    case 16: return sparse_kernel<MIXED, GEN, FORM, 16>;
    case 32: return sparse_kernel<MIXED, GEN, FORM, 32>;
    case 64:
        if constexpr (MIXED) return sparse_kernel<MIXED, GEN, FORM, 64>;
        else return nullptr;
    default: return nullptr;
    }
}

SparseKernel pick(bool mixed, int gen, int form, int ring)
{
    const bool sob = gen != kGenRichtmyer;
    if (mixed)
        return sob ? pick_ring<true, kGenSobol, kPhiTail>(ring) : pick_ring<true, kGenRichtmyer, kPhiTail>(ring);
    if (form == kPhiTail)
        return sob ? pick_ring<false, kGenSobol, kPhiTail>(ring) : pick_ring<false, kGenRichtmyer, kPhiTail>(ring);
    return sob ? pick_ring<false, kGenSobol, kPhiErf>(ring) : pick_ring<false, kGenRichtmyer, kPhiErf>(ring);
}

// Resident blocks of the ring kernel on the whole device (LDS and registers
// permitting); the workspace kernel keeps the plan's grid.
hipError_t sparse_grid(bool mixed, int gen, int form, int ring, int ncu, int* grid)
{
    const SparseKernel fn = pick(mixed, gen, mixed ? kPhiTail : form, ring);
    if (fn == nullptr) return hipErrorInvalidValue;
    int per_cu = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fn, kTileS, 0);
    if (e != hipSuccess) return e;
    *grid = ncu * std::max(per_cu, 1);
    return hipSuccess;
}

hipError_t launch_sparse(const SparseArgs& A, bool mixed, int gen, int ring, int grid, hipStream_t s)
{
    const SparseKernel fn = pick(mixed, gen, mixed ? kPhiTail : A.form, ring);
    if (fn == nullptr) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(kTileS), 0, s, A, DBG_POINT_ARGS(A.pts));
    return hipGetLastError();
}

}  // namespace dbg
