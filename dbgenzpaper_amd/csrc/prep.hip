// Preparation of a factor for the integration kernels:
//   * truncate_matrixU!  (src/qmc.jl:133-142) on the upper triangle,
//   * the packed row layout of the strict upper factor read by the dense
//     kernels, with the per-target bounds and 1/(U_jj sqrt 2),
//   * the first-dimension interval of qmc_pnorm! (src/qmc.jl:800-809),
//   * the CSC of triu(U, 1) of the sparse constructor (src/qmc.jl:272-279).
#include "dbg_device.h"

namespace dbg {

__global__ void truncate_kernel(double* U, int64_t lda, int n, double min_val)
{
    const int j = blockIdx.y;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= j && i < n; i += gridDim.x * blockDim.x) {
        double& v = U[(size_t)i + (size_t)j * lda];
        if (fabs(v) < min_val) v = 0.0;
    }
}

hipError_t truncate_upper(double* U, int64_t lda, int n, double min_val, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    dim3 g((n + 255) / 256, n), b(256);
    hipLaunchKernelGGL(truncate_kernel, g, b, 0, s, U, lda, n, min_val);
    return hipGetLastError();
}

// Uk[l * ldu + j] = U[l, j] for j > l (rows l = 0 .. n-2), zero elsewhere.
template <typename T>
__global__ void pack_rows_kernel(const double* __restrict__ U, int64_t lda, int n, T* Uk, int64_t ldu)
{
    const int l = blockIdx.y;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < ldu; j += (int64_t)gridDim.x * blockDim.x) {
        double v = (j > l && j < n) ? U[(size_t)l + (size_t)j * lda] : 0.0;
        Uk[(size_t)l * ldu + j] = (T)v;
    }
}

template <typename T>
__global__ void pack_vec_kernel(const double* __restrict__ U, int64_t lda, int n, const double* a,
                                const double* b, T* ak, T* bk, T* invk, double* first, int phi_form,
                                bool mixed)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) {
        const double ujj = U[(size_t)j + (size_t)j * lda];
        ak[j] = (T)a[j];
        bk[j] = (T)b[j];
        // inv = 1 / (U_jj * sqrt 2) in the working precision (src/qmc.jl:407)
        if (mixed)
            invk[j] = (T)(1.0 / (ujj * kSqrt2));
        else
            invk[j] = (T)(1.0 / __dmul_rn(ujj, kSqrt2));
    }
    if (j == 0) {
        const double u11 = U[0];
        first_from_args(a[0] / __dmul_rn(u11, kSqrt2), b[0] / __dmul_rn(u11, kSqrt2), phi_form, first);
    }
}

hipError_t pack_dense(const double* U, int64_t lda, int n, const double* a, const double* b,
                      bool mixed, void* Uk, int64_t ldu, void* ak, void* bk, void* invk,
                      double* first, int phi_form, hipStream_t s)
{
    if (Uk != nullptr && n > 1) {
        dim3 g((unsigned)((ldu + 255) / 256), n - 1), bl(256);
        if (mixed)
            hipLaunchKernelGGL(pack_rows_kernel<float>, g, bl, 0, s, U, lda, n, (float*)Uk, ldu);
        else
            hipLaunchKernelGGL(pack_rows_kernel<double>, g, bl, 0, s, U, lda, n, (double*)Uk, ldu);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    dim3 g2((n + 255) / 256), b2(256);
    if (mixed)
        hipLaunchKernelGGL(pack_vec_kernel<float>, g2, b2, 0, s, U, lda, n, a, b, (float*)ak,
                           (float*)bk, (float*)invk, first, phi_form, true);
    else
        hipLaunchKernelGGL(pack_vec_kernel<double>, g2, b2, 0, s, U, lda, n, a, b, (double*)ak,
                           (double*)bk, (double*)invk, first, phi_form, false);
    return hipGetLastError();
}

__global__ void convert_f32_kernel(const double* __restrict__ in, size_t count, float* out)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

hipError_t convert_f32(const double* in, size_t count, float* out, hipStream_t s)
{
    const size_t blocks = std::min<size_t>((count + 255) / 256, 65536);
    hipLaunchKernelGGL(convert_f32_kernel, dim3((unsigned)std::max<size_t>(blocks, 1)), dim3(256), 0, s,
                       in, count, out);
    return hipGetLastError();
}

// --- CSC of triu(U, 1): rows ascending within each column ------------------
__global__ void csc_count_kernel(const double* __restrict__ U, int64_t lda, int n, int64_t* cnt)
{
    const int j = blockIdx.x;
    __shared__ int64_t s_cnt[4];
    int64_t c = 0;
    for (int i = threadIdx.x; i < j; i += blockDim.x) c += (U[(size_t)i + (size_t)j * lda] != 0.0);
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0) s_cnt[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) cnt[j] = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
}

hipError_t csc_count(const double* U, int64_t lda, int n, int64_t* colcount, hipStream_t s)
{
    hipLaunchKernelGGL(csc_count_kernel, dim3(n), dim3(256), 0, s, U, lda, n, colcount);
    return hipGetLastError();
}

// exclusive scan of n counts into n+1 pointers (single block, sequential
// chunks; n is a dimension, at most tens of thousands)
__global__ void scan_kernel(const int64_t* in, int64_t* out, int n)
{
    __shared__ int64_t part[256];
    const int tid = threadIdx.x;
    const int chunk = (n + 255) / 256;
    const int lo = tid * chunk, hi = min(n, lo + chunk);
    int64_t s = 0;
    for (int i = lo; i < hi; ++i) s += in[i];
    part[tid] = s;
    __syncthreads();
    if (tid == 0) {
        int64_t run = 0;
        for (int t = 0; t < 256; ++t) {
            const int64_t v = part[t];
            part[t] = run;
            run += v;
        }
        out[n] = run;
    }
    __syncthreads();
    int64_t run = part[tid];
    for (int i = lo; i < hi; ++i) {
        out[i] = run;
        run += in[i];
    }
}

hipError_t exclusive_scan_i64(const int64_t* in, int64_t* out, int n, hipStream_t s)
{
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

// One wavefront per column: ballot-compacted, rows in ascending order.
template <typename T>
__global__ void csc_fill_kernel(const double* __restrict__ U, int64_t lda, int n,
                                const int64_t* __restrict__ colptr, int32_t* rowval, T* nzval)
{
    const int j = blockIdx.x;
    const int lane = threadIdx.x;
    int64_t pos = colptr[j];
    for (int i0 = 0; i0 < j; i0 += 64) {
        const int i = i0 + lane;
        const double v = (i < j) ? U[(size_t)i + (size_t)j * lda] : 0.0;
        const bool nz = (i < j) && (v != 0.0);
        const uint64_t ball = __ballot(nz);
        if (nz) {
            const int before = __popcll(ball & ((1ull << lane) - 1ull));
            rowval[pos + before] = i;
            nzval[pos + before] = (T)v;
        }
        pos += __popcll(ball);
    }
}

hipError_t csc_fill(const double* U, int64_t lda, int n, const int64_t* colptr, bool mixed,
                    int32_t* rowval, void* nzval, hipStream_t s)
{
    if (mixed)
        hipLaunchKernelGGL(csc_fill_kernel<float>, dim3(n), dim3(64), 0, s, U, lda, n, colptr,
                           rowval, (float*)nzval);
    else
        hipLaunchKernelGGL(csc_fill_kernel<double>, dim3(n), dim3(64), 0, s, U, lda, n, colptr,
                           rowval, (double*)nzval);
    return hipGetLastError();
}

// Reach of the pattern: the largest g - r over the rows r that column g + 1
// reads when dimension g + 1 is conditioned on y_0..y_g (rows ascend, so a
// column's first row is its farthest); -1 when triu(U, 1) is empty.  One
// block: a max over n columns.
__global__ void csc_reach_kernel(const int64_t* __restrict__ colptr, const int32_t* __restrict__ rowval,
                                 int n, int* out)
{
    __shared__ int s_max[4];
    int r = -1;
    for (int j = threadIdx.x; j < n; j += blockDim.x)
        if (colptr[j + 1] > colptr[j]) r = max(r, j - 1 - rowval[colptr[j]]);
    for (int off = 32; off > 0; off >>= 1) r = max(r, __shfl_xor(r, off, 64));
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3]));
}

hipError_t csc_reach(const int64_t* colptr, const int32_t* rowval, int n, int* out, hipStream_t s)
{
    hipLaunchKernelGGL(csc_reach_kernel, dim3(1), dim3(256), 0, s, colptr, rowval, n, out);
    return hipGetLastError();
}

}  // namespace dbg
