// Covariance generators of the reference's simulation scripts
// (scripts/simulations/simulations_functions.jl), built on the device so
// that a job's inputs need not come from the host: element-wise fills for
// the closed-form matrices, rocBLAS dsyrk for the low-rank-plus-diagonal
// ones, then an optional standardization to a correlation matrix.
#include <rocblas/rocblas.h>

#include "dbg_device.h"

namespace dbg {

namespace {

#define SS(i, j) S[(size_t)(i) + (size_t)(j) * (size_t)ld]

// kind: kCovFixedDense ones(n, n) + p0 I (:21-24); kCovFixedSparse the
// banded matrix of fixed_sparse(n, p0) (:33-41: p0 on the diagonal, 1 within
// p0 of it); kCovMatern p2 * M_nu(sqrt(2 nu) |i - j| / p1), nu = p0 in
// {0.5, 1.5, 2.5} (:56-79; the half-integer closed forms of
// 2^(1-nu)/Gamma(nu) z^nu K_nu(z))
__global__ void cov_fill_kernel(int kind, int n, double* __restrict__ S, int64_t ld, double p0, double p1,
                                double p2)
{
    const int j = blockIdx.y;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int d = i > j ? i - j : j - i;
        double v;
        if (kind == kCovFixedDense) {
            v = i == j ? 1.0 + p0 : 1.0;
        } else if (kind == kCovFixedSparse) {
            v = i == j ? p0 : (d <= (int)p0 ? 1.0 : 0.0);
        } else {
            if (d == 0) {
                v = p2;
            } else {
                const double z = sqrt(2.0 * p0) * (double)d / p1;
                const double e = exp(-z);
                if (p0 == 0.5) v = p2 * e;
                else if (p0 == 1.5) v = p2 * (1.0 + z) * e;
                else v = p2 * (1.0 + z + z * z / 3.0) * e;
            }
        }
        SS(i, j) = v;
    }
}

// the upper triangle of B B' (dsyrk) -> the full matrix, psi on the diagonal
__global__ void cov_lowrank_finish_kernel(int n, double* __restrict__ S, int64_t ld, const double* __restrict__ psi)
{
    const int j = blockIdx.y;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (i > j) SS(i, j) = SS(j, i);
        else if (i == j) SS(i, i) += psi[i];
    }
}

__global__ void cov_diag_sqrt_kernel(int n, const double* __restrict__ S, int64_t ld, double* __restrict__ s)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) s[i] = sqrt(SS(i, i));
}

// S_ij / (s_i s_j), exact ones on the diagonal (the order of the host
// generators' standardization)
__global__ void cov_standardize_kernel(int n, double* __restrict__ S, int64_t ld, const double* __restrict__ s)
{
    const int j = blockIdx.y;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        SS(i, j) = i == j ? 1.0 : SS(i, j) / (s[i] * s[j]);
}

#undef SS

}  // namespace

hipError_t cov_fill(int kind, int n, double* S, int64_t ld, double p0, double p1, double p2, hipStream_t s)
{
    const dim3 g((n + 255) / 256, n);
    hipLaunchKernelGGL(cov_fill_kernel, g, dim3(256), 0, s, kind, n, S, ld, p0, p1, p2);
    return hipGetLastError();
}

int cov_lowrank(int n, int k, const double* B, const double* psi, double* S, int64_t ld, hipStream_t s)
{
    rocblas_handle h = (rocblas_handle)blas_handle_void(s);
    if (!h) return -1;
    const double one = 1.0, zero = 0.0;
    if (rocblas_dsyrk(h, rocblas_fill_upper, rocblas_operation_none, n, k, &one, B, n, &zero, S, ld) !=
        rocblas_status_success)
        return -1;
    const dim3 g((n + 255) / 256, n);
    hipLaunchKernelGGL(cov_lowrank_finish_kernel, g, dim3(256), 0, s, n, S, ld, psi);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

hipError_t cov_standardize(int n, double* S, int64_t ld, double* scratch, hipStream_t s)
{
    hipLaunchKernelGGL(cov_diag_sqrt_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, S, ld, scratch);
    const dim3 g((n + 255) / 256, n);
    hipLaunchKernelGGL(cov_standardize_kernel, g, dim3(256), 0, s, n, S, ld, scratch);
    return hipGetLastError();
}

}  // namespace dbg
