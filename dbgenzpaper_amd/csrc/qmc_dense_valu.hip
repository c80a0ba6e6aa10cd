// Dense integration kernel, vector-ALU form: one sample per lane.
//
// This is the qmc_loop! hot loop of src/qmc.jl:315-444 (and the erfc variant
// :485-608) with the point generation of src/qmc_generators.jl:137-204 fused
// in.  A block of 256 lanes owns 256 points of one shift.  The sampled
// dimensions are swept in blocks of kW: the conditional means s of the kW
// target dimensions of a block are first accumulated from all earlier
// dimensions (the outer GEMM of :354, here a register-blocked dot with the
// packed U row broadcast from the scalar cache), then the block is walked one
// dimension at a time with the in-block triangular update (:390-403) done in
// registers with compile-time indices.  Each sample's y values are written to
// a per-block workspace slot and read back by the later blocks.
//
// Serves every dimension; it is the only path for small d, where the work is
// bound by the special functions rather than the contraction.
#include <type_traits>

#include "dbg_device.h"

namespace dbg {

constexpr int kTileV = 256;
constexpr int kWV = 16;

int dense_valu_tile() { return kTileV; }
int dense_valu_width() { return kWV; }

template <bool MIXED, int GEN, int FORM>
__global__ __launch_bounds__(kTileV) void dense_valu_kernel(DenseArgs A)
{
    using T = typename std::conditional<MIXED, float, double>::type;
    using State = typename std::conditional<MIXED, StateMixed, StateF64>::type;
    const int tid = threadIdx.x;
    const int nd = A.n - 1;
    const T* __restrict__ Uk = (const T*)A.Uk;
    const T* __restrict__ ak = (const T*)A.ak;
    const T* __restrict__ bk = (const T*)A.bk;
    const T* __restrict__ invk = (const T*)A.invk;
    T* __restrict__ Ys = (T*)A.Y + (size_t)blockIdx.x * (size_t)A.y_slot;

    for (int64_t tile = blockIdx.x; tile < A.n_tiles; tile += gridDim.x) {
        const int k = (int)(tile / A.tiles_per_rep);
        const int64_t jt = tile - (int64_t)k * A.tiles_per_rep;
        const int64_t j = jt * kTileV + tid + 1;  // 1-based point index of the round
        const bool valid = j <= A.m;

        State st;
        Weight<T> w;
        if constexpr (MIXED) {
            st.ea = A.ea1;
            st.eb = A.eb1;
            st.dc = A.d1;
            st.mode = A.mode1;
            w.init(A.d1);
        } else {
            st.c = A.c1;
            st.dc = A.dc1;
            w.init(A.dc1);
        }

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
            // walk the block
#pragma unroll
            for (int cc = 0; cc < kWV; ++cc) {
                const int g = g0 + cc;
                if (g < nd) {
                    const double u = point_u<GEN>(A.pts, g, k, j);
                    T y;
                    if constexpr (MIXED)
                        y = sample_mixed(st, u);
                    else
                        y = sample_f64(st, u);
                    Ys[(size_t)g * kTileV + tid] = y;
                    const T* __restrict__ ur = Uk + (size_t)g * A.ldu + g0 + 1;
#pragma unroll
                    for (int c2 = cc; c2 < kWV; ++c2) s[c2] = fma(y, ur[c2], s[c2]);
                    if constexpr (MIXED)
                        phi_mixed(st, s[cc], ak[g + 1], bk[g + 1], invk[g + 1]);
                    else
                        phi_f64<FORM>(st, s[cc], ak[g + 1], bk[g + 1], invk[g + 1]);
                    w.mul(st.dc);
                }
            }
        }
        const double pm = valid ? (double)w.pm : 0.0;
        double S;
        int E;
        block_scaled_sum<kTileV>(pm, w.pe, S, E);
        if (tid == 0) {
            A.part_m[tile] = S;
            A.part_e[tile] = E;
        }
    }
}

hipError_t launch_dense_valu(const DenseArgs& A, bool mixed, int gen, int form, int grid,
                             hipStream_t s)
{
    dim3 g(grid), b(kTileV);
    if (mixed) {
        if (gen == kGenRichtmyer)
            hipLaunchKernelGGL((dense_valu_kernel<true, kGenRichtmyer, kPhiErf>), g, b, 0, s, A);
        else
            hipLaunchKernelGGL((dense_valu_kernel<true, kGenSobol, kPhiErf>), g, b, 0, s, A);
    } else if (gen == kGenRichtmyer) {
        if (form == kPhiErf)
            hipLaunchKernelGGL((dense_valu_kernel<false, kGenRichtmyer, kPhiErf>), g, b, 0, s, A);
        else
            hipLaunchKernelGGL((dense_valu_kernel<false, kGenRichtmyer, kPhiErfc>), g, b, 0, s, A);
    } else {
        if (form == kPhiErf)
            hipLaunchKernelGGL((dense_valu_kernel<false, kGenSobol, kPhiErf>), g, b, 0, s, A);
        else
            hipLaunchKernelGGL((dense_valu_kernel<false, kGenSobol, kPhiErfc>), g, b, 0, s, A);
    }
    return hipGetLastError();
}

}  // namespace dbg
