// Rank-deficient Sigma (no reference counterpart: QMCData throws for it,
// src/qmc.jl:195-197).  A dense plan whose Genz factorization stops short of
// full rank is factored again with the row-max pivoting of cholesky_rowmax!
// (src/cholesky.jl:266-432) and a rank tolerance, the rows 0 .. r-1 of the
// dependent columns are completed from Sigma (U_r^T X = C_p[0:r, r:n], rocBLAS
// dtrsm), and adapt_low_rank! (src/cholesky.jl:20-123) groups the columns
// into blocks, one per latent variable i: its own column first (U[i, c] =
// U_ii), then the dependent variables whose last coefficient is on latent i
// (scaled to about 1, bounds exchanged for a negative one).  The integration
// then walks the r latent variables: the interval of latent i is the
// intersection over its block of
//     [(a_c - s_c) inv_c, (b_c - s_c) inv_c],  inv_c = 1 / (U[i, c] sqrt 2),
//     s_c = sum_{l < i} U[l, c] y_l,
// in erf-argument units (empty: dc_i = 0), and y_i is drawn from it with
// point dimension i, as in the dense loop, which this is when every block has
// one column.  Columns beg[r] .. n-1 (no coefficient, zero variance) multiply
// the estimate by [a_c <= 0 <= b_c].  FP64 throughout, one sample per lane,
// the y of a sample in a per-workgroup slot of (r - 1) x 256 values.
#include "dbg_device.h"

namespace dbg {

constexpr int kTileL = 256;

int lowrank_tile() { return kTileL; }

// U[i, j] for i < r <= j from Sigma: C_p[i, j] = Sigma[p_i, p_j] (1 / s_{p_j})
// / s_{p_i} (the standardization of src/cholesky.jl:285-300), the right-hand
// side of U_r^T X = C_p[0:r, r:n]; rows >= r and the strict lower triangle of
// the first r columns zeroed.  Grid (x, n).
__global__ void lowrank_fill_kernel(double* __restrict__ U, int n, int r, const double* __restrict__ Sig,
                                    const int32_t* __restrict__ perm, const double* __restrict__ scale)
{
    const int j = blockIdx.y;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double& u = U[(size_t)i + (size_t)j * n];
        if (i >= r || (j < r && i > j)) {
            u = 0.0;
        } else if (j >= r) {
            const int pi = perm[i], pj = perm[j];
            u = Sig[(size_t)pi + (size_t)pj * n] * ((1.0 / scale[pj]) / scale[pi]);
        }
    }
}

hipError_t lowrank_fill(double* U, int n, int r, const double* Sig, const int32_t* perm, const double* scale,
                        hipStream_t s)
{
    hipLaunchKernelGGL(lowrank_fill_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, U, n, r, Sig, perm, scale);
    return hipGetLastError();
}

// ap[i] = a[perm[i]], bp likewise (cholesky_rowmax! leaves the standardized
// bounds in the original order)
__global__ void lowrank_bounds_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                      const int32_t* __restrict__ perm, int n, double* __restrict__ ap,
                                      double* __restrict__ bp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        ap[i] = a[perm[i]];
        bp[i] = b[perm[i]];
    }
}

hipError_t lowrank_bounds(const double* a, const double* b, const int32_t* perm, int n, double* ap, double* bp,
                          hipStream_t s)
{
    hipLaunchKernelGGL(lowrank_bounds_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a, b, perm, n, ap, bp);
    return hipGetLastError();
}

// inv[c] = 1 / (U[i, c] sqrt 2) for the columns c of block i (one thread per
// block); thread 0 also forms latent 0's interval, the intersection over
// block 0 of [a_c, b_c] / (U[0, c] sqrt 2), with the zero-variance columns'
// factor, in the first-dimension layout of pack_dense.
__global__ void lowrank_prep_kernel(const double* __restrict__ U, int n, const int32_t* __restrict__ beg, int r,
                                    const double* __restrict__ a, const double* __restrict__ b,
                                    double* __restrict__ inv, double* __restrict__ first, int phi_form)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < r)
        for (int c = beg[i]; c < beg[i + 1]; ++c) inv[c] = 1.0 / __dmul_rn(U[(size_t)i + (size_t)c * n], kSqrt2);
    if (i == 0) {
        double za = -INFINITY, zb = INFINITY;
        for (int c = beg[0]; c < beg[1]; ++c) {
            const double u = __dmul_rn(U[(size_t)c * n], kSqrt2);
            za = fmax(za, a[c] / u);
            zb = fmin(zb, b[c] / u);
        }
        if (!(za <= zb)) zb = za;
        first_from_args(za, zb, phi_form, first);
        for (int c = beg[r]; c < n; ++c) {
            if (!(a[c] <= 0.0 && 0.0 <= b[c])) {
                first[1] = 0.0;
                first[4] = 0.0;
            }
        }
    }
}

hipError_t lowrank_prep(const double* U, int n, const int32_t* beg, int r, const double* a, const double* b,
                        double* inv, double* first, int phi_form, hipStream_t s)
{
    hipLaunchKernelGGL(lowrank_prep_kernel, dim3((r + 255) / 256), dim3(256), 0, s, U, n, beg, r, a, b, inv, first,
                       phi_form);
    return hipGetLastError();
}

// The integration (one round over the plan's block of points).  Ucol: the
// reordered factor (column-major, ldc), ak / bk / invk per column, lr_beg the
// r + 1 block starts.  The factor entries are uniform across the lanes
// (scalar loads); the sample's y go to its workgroup's slot row by row.
template <int GEN, int FORM>
__global__ __launch_bounds__(kTileL) void lowrank_kernel(DenseArgs A, const double* __restrict__ U,
                                                         const double* __restrict__ ak,
                                                         const double* __restrict__ bk,
                                                         const double* __restrict__ invk,
                                                         const int32_t* __restrict__ beg, double* __restrict__ Yall,
                                                         double* __restrict__ part_m, int* __restrict__ part_e,
                                                         DBG_POINT_PARAMS)
{
    using PT = Path<false, FORM>;
    const PointArgs pts = DBG_WITH_TABLES(A.pts);
    const int tid = threadIdx.x;
    const int r = A.lr_rank;
    const int64_t ldc = A.ldc;
    double* __restrict__ Ys = Yall + (size_t)blockIdx.x * (size_t)A.y_slot;

    for (int64_t tile = blockIdx.x; tile < A.n_tiles; tile += gridDim.x) {
        const int k = (int)(tile / A.tiles_per_rep);
        const int64_t jt = tile - (int64_t)k * A.tiles_per_rep;
        const int64_t j = jt * kTileL + tid + 1;  // 1-based point index of the round
        const bool valid = j <= A.m;

        typename PT::State st;
        Weight<double> w;
        PT::init(st, w, A.first);
        for (int g = 0; g + 1 < r; ++g) {
            const PointOp op = point_op<false, GEN>(pts, g, k);
            Ys[(size_t)g * kTileL + tid] = path_draw<false, GEN, FORM>(pts, op, st, g, j);
            const int i = g + 1;
            double za = -INFINITY, zb = INFINITY;
            const int c1 = beg[i + 1];
            for (int c = beg[i]; c < c1; ++c) {
                const double* __restrict__ uc = U + (size_t)c * ldc;
                double sc = 0.0;
                for (int l = 0; l < i; ++l) sc = fma(Ys[(size_t)l * kTileL + tid], uc[l], sc);
                za = fmax(za, __dmul_rn(__dsub_rn(ak[c], sc), invk[c]));
                zb = fmin(zb, __dmul_rn(__dsub_rn(bk[c], sc), invk[c]));
            }
            if (!(za <= zb)) zb = za;
            PT::phi_z(st, za, zb);
            PT::weigh(w, st);
        }
        const double pm = valid ? w.pm : 0.0;
        double S;
        int E;
        block_scaled_sum<kTileL>(pm, w.pe, S, E);
        if (tid == 0) {
            part_m[tile] = S;
            part_e[tile] = E;
        }
    }
}

template <int GEN, int FORM>
static void launch_l(const DenseArgs& A, int grid, hipStream_t s)
{
    hipLaunchKernelGGL((lowrank_kernel<GEN, FORM>), dim3(grid), dim3(kTileL), 0, s, A, (const double*)A.Ucol,
                       (const double*)A.ak, (const double*)A.bk, (const double*)A.invk, A.lr_beg, (double*)A.Y,
                       A.part_m, A.part_e, DBG_POINT_ARGS(A.pts));
}

hipError_t launch_lowrank(const DenseArgs& A, int gen, int form, int grid, hipStream_t s)
{
    if (gen == kGenRichtmyer) {
        if (form == kPhiErf) launch_l<kGenRichtmyer, kPhiErf>(A, grid, s);
        else if (form == kPhiErfc) launch_l<kGenRichtmyer, kPhiErfc>(A, grid, s);
        else launch_l<kGenRichtmyer, kPhiTail>(A, grid, s);
    } else {
        if (form == kPhiErf) launch_l<kGenSobol, kPhiErf>(A, grid, s);
        else if (form == kPhiErfc) launch_l<kGenSobol, kPhiErfc>(A, grid, s);
        else launch_l<kGenSobol, kPhiTail>(A, grid, s);
    }
    return hipGetLastError();
}

}  // namespace dbg
