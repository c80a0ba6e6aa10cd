// Sparse / banded integration kernel: qmc_loop!(::QMCDataSparse)
// (src/qmc.jl:618-718) and SparseQMC.qmc_loop! (src/qmc_sparse.jl:159-262).
//
// One sample per lane, 256 points of one shift per block.  For each sampled
// dimension the conditional mean of the next dimension is the dot of the
// stored column of triu(U, 1) with the sample's earlier y values, in CSC row
// order.  Row indices and values are the same for every lane (scalar loads);
// the y values come from the block's workspace slot, written moments earlier
// by the same lane, so for a banded factor they are served from the L1/L2.
// No contraction here: the kernel is bound by the special functions.
#include <type_traits>

#include "dbg_device.h"

namespace dbg {

constexpr int kTileS = 256;
int sparse_tile() { return kTileS; }

template <bool MIXED, int GEN, int FORM>
__global__ __launch_bounds__(kTileS) void sparse_kernel(SparseArgs A, DBG_POINT_PARAMS)
{
    const PointArgs pts = DBG_WITH_TABLES(A.pts);
    using T = typename std::conditional<MIXED, float, double>::type;
    using PT = Path<MIXED, FORM>;
    const int tid = threadIdx.x;
    const int nd = A.n - 1;
    const int64_t* __restrict__ colptr = A.colptr;
    const int32_t* __restrict__ rowval = A.rowval;
    const T* __restrict__ nzval = (const T*)A.nzval;
    const T* __restrict__ ak = (const T*)A.ak;
    const T* __restrict__ bk = (const T*)A.bk;
    const T* __restrict__ invk = (const T*)A.invk;
    T* __restrict__ Ys = (T*)A.Y + (size_t)blockIdx.x * (size_t)A.y_slot;

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
                y = tail_sample_f(st, wf, wc);
            } else {
                const double u = point_u<GEN>(pts, g, k, j);
                y = A.tent ? path_sample<MIXED, FORM>(st, u) : PT::sample_w(st, u);
            }
            Ys[(size_t)g * kTileS + tid] = y;
            T sv = T(0);
            const int64_t p1 = colptr[g + 2];
            for (int64_t p = colptr[g + 1]; p < p1; ++p)
                sv = fma(Ys[(size_t)rowval[p] * kTileS + tid], nzval[p], sv);
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

hipError_t launch_sparse(const SparseArgs& A, bool mixed, int gen, int grid, hipStream_t s)
{
    dim3 g(grid), b(kTileS);
    if (mixed) {
        if (gen == kGenRichtmyer)
            hipLaunchKernelGGL((sparse_kernel<true, kGenRichtmyer, kPhiTail>), g, b, 0, s, A, DBG_POINT_ARGS(A.pts));
        else
            hipLaunchKernelGGL((sparse_kernel<true, kGenSobol, kPhiTail>), g, b, 0, s, A, DBG_POINT_ARGS(A.pts));
    } else if (A.form == kPhiTail) {
        if (gen == kGenRichtmyer)
            hipLaunchKernelGGL((sparse_kernel<false, kGenRichtmyer, kPhiTail>), g, b, 0, s, A, DBG_POINT_ARGS(A.pts));
        else
            hipLaunchKernelGGL((sparse_kernel<false, kGenSobol, kPhiTail>), g, b, 0, s, A, DBG_POINT_ARGS(A.pts));
    } else {
        if (gen == kGenRichtmyer)
            hipLaunchKernelGGL((sparse_kernel<false, kGenRichtmyer, kPhiErf>), g, b, 0, s, A, DBG_POINT_ARGS(A.pts));
        else
            hipLaunchKernelGGL((sparse_kernel<false, kGenSobol, kPhiErf>), g, b, 0, s, A, DBG_POINT_ARGS(A.pts));
    }
    return hipGetLastError();
}

}  // namespace dbg
