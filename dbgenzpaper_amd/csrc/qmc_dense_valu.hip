// Dense integration kernel, vector-ALU form: one sample per lane.
//
// This is the qmc_loop! hot loop of src/qmc.jl:315-444 (and the erfc variant
// :485-608) with the point generation of src/qmc_generators.jl:137-204 fused
// in.  A block of 256 lanes owns 256 points of one shift.  The sampled
// dimensions are swept in blocks of 16: the conditional means of the block's
// 16 targets are first accumulated from all earlier dimensions (the outer mul!
// of :354, here a register-blocked dot with the packed U row broadcast from
// the scalar cache), handed to LDS, and the block is walked one dimension at a
// time by walk_block.  Each sample's y values are written to a per-block
// workspace slot and read back by the later blocks.
//
// Used for small d, where the work is bound by the special functions rather
// than by the contraction.
#include <type_traits>

#include "dbg_device.h"

namespace dbg {

constexpr int kTileV = 256;
constexpr int kWV = 16;
constexpr int kSXV = kWV + 4;

int dense_valu_tile() { return kTileV; }
int dense_valu_width() { return kWV; }

template <bool MIXED, int GEN, int FORM, typename T = typename std::conditional<MIXED, float, double>::type>
__global__ __launch_bounds__(kTileV) void dense_valu_kernel(DenseArgs A, const T* __restrict__ Uk,
                                                            const T* __restrict__ Ucol,
                                                            const T* __restrict__ ak,
                                                            const T* __restrict__ bk,
                                                            const T* __restrict__ invk,
                                                            T* __restrict__ Yall,
                                                            double* __restrict__ part_m,
                                                            int* __restrict__ part_e, DBG_POINT_PARAMS)
{
    const PointArgs pts = DBG_WITH_TABLES(A.pts);
    __shared__ T s_x[kTileV * kSXV];
    __shared__ T s_y[kWV * kTileV];
    const int tid = threadIdx.x;
    const int nd = A.n - 1;
    T* __restrict__ Ys = Yall + (size_t)blockIdx.x * (size_t)A.y_slot;

    for (int64_t it = blockIdx.x; it < A.n_tiles; it += gridDim.x) {
        const int64_t tile = A.tile_list ? A.tile_list[it] : it;
        if (tile < 0) continue;  // uniform: one entry per workgroup
        const int k = (int)(tile / A.tiles_per_rep);
        const int64_t jt = tile - (int64_t)k * A.tiles_per_rep;
        const int64_t j = jt * kTileV + tid + 1;  // 1-based point index of the round
        const bool valid = j <= A.m;

        typename Path<MIXED, FORM>::State st;
        Weight<T> w;
        Path<MIXED, FORM>::init(st, w, A.first);

        for (int g0 = 0; g0 < nd; g0 += kWV) {
            T s[kWV];
#pragma unroll
            for (int c = 0; c < kWV; ++c) s[c] = T(0);
            // contributions of all earlier dimensions to the block's targets
            for (int l = 0; l < g0; ++l) {
                const T y = Ys[(size_t)l * kTileV + tid];
                const T* __restrict__ ur = Uk + (size_t)l * A.ldu + g0 + 1;
#pragma unroll
                for (int c = 0; c < kWV; ++c) s[c] = fma(y, ur[c], s[c]);
            }
#pragma unroll
            for (int c = 0; c < kWV; ++c) s_x[tid * kSXV + c] = s[c];
            walk_block<MIXED, GEN, FORM, kTileV, kSXV, kWV>(pts, g0, min(kWV, nd - g0), k, j, tid, st, w,
                                                          s_x, s_y, Ys, kTileV, Ucol, A.ldc, ak, bk, invk);
        }
        const double pm = valid ? (double)w.pm : 0.0;
        double S;
        int E;
        block_scaled_sum<kTileV>(pm, w.pe, S, E);
        if (tid == 0) {
            part_m[tile] = S;
            part_e[tile] = E;
        }
    }
}

template <bool MIXED, int GEN, int FORM>
static void launch_v(const DenseArgs& A, int grid, hipStream_t s)
{
    using T = typename std::conditional<MIXED, float, double>::type;
    hipLaunchKernelGGL((dense_valu_kernel<MIXED, GEN, FORM>), dim3(grid), dim3(kTileV), 0, s, A,
                       (const T*)A.Uk, (const T*)A.Ucol, (const T*)A.ak, (const T*)A.bk,
                       (const T*)A.invk, (T*)A.Y, A.part_m, A.part_e, DBG_POINT_ARGS(A.pts));
}

hipError_t launch_dense_valu(const DenseArgs& A, bool mixed, int gen, int form, int grid,
                             hipStream_t s)
{
    if (mixed) {
        // a lower orthant takes the walk that forms only the upper bound's tail mass
        if (gen == kGenRichtmyer) {
            if (A.a_inf) launch_v<true, kGenRichtmyer, kPhiTailLower>(A, grid, s);
            else launch_v<true, kGenRichtmyer, kPhiTail>(A, grid, s);
        } else {
            if (A.a_inf) launch_v<true, kGenSobol, kPhiTailLower>(A, grid, s);
            else launch_v<true, kGenSobol, kPhiTail>(A, grid, s);
        }
    } else if (gen == kGenRichtmyer) {
        if (form == kPhiErf)
            launch_v<false, kGenRichtmyer, kPhiErf>(A, grid, s);
        else if (form == kPhiErfc)
            launch_v<false, kGenRichtmyer, kPhiErfc>(A, grid, s);
        else
            launch_v<false, kGenRichtmyer, kPhiTail>(A, grid, s);
    } else {
        if (form == kPhiErf)
            launch_v<false, kGenSobol, kPhiErf>(A, grid, s);
        else if (form == kPhiErfc)
            launch_v<false, kGenSobol, kPhiErfc>(A, grid, s);
        else
            launch_v<false, kGenSobol, kPhiTail>(A, grid, s);
    }
    return hipGetLastError();
}

}  // namespace dbg
