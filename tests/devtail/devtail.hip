// Test support: evaluates the FP32 tail functions of the mixed-precision
// walk (tailq.h) on the GPU, so that the tests can check the device code
// (v_rcp_f32, v_exp_f32, v_log_f32 in place of the host's library calls)
// against high-precision values.  Not part of libdbgenz.so.
#include <hip/hip_runtime.h>

#include "tailq.h"

namespace {

__global__ void devtail_kernel(const float* __restrict__ z, const float* __restrict__ t, const int* __restrict__ tk,
                               int n, float* __restrict__ m, int* __restrict__ k, float* __restrict__ y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int kk = 0;
    m[i] = dbg::tail_mass_f32(z[i], kk);
    k[i] = kk;
    y[i] = dbg::tail_quantile_f32(t[i], tk[i]);
}

}  // namespace

// m[i] 2^-k[i] = erfc(z[i])/2 and y[i] = Phi^{-1}(t[i] 2^-tk[i]) for i < n
// (host arrays).  Returns 0 or the HIP error code.
extern "C" int devtail_eval(const float* z, const float* t, const int* tk, int n, float* m, int* k, float* y)
{
    if (n <= 0) return 0;
    const size_t fb = (size_t)n * sizeof(float), ib = (size_t)n * sizeof(int);
    float *dz = nullptr, *dt = nullptr, *dm = nullptr, *dy = nullptr;
    int *dtk = nullptr, *dk = nullptr;
    hipError_t e = hipMalloc((void**)&dz, fb);
    if (e == hipSuccess) e = hipMalloc((void**)&dt, fb);
    if (e == hipSuccess) e = hipMalloc((void**)&dm, fb);
    if (e == hipSuccess) e = hipMalloc((void**)&dy, fb);
    if (e == hipSuccess) e = hipMalloc((void**)&dtk, ib);
    if (e == hipSuccess) e = hipMalloc((void**)&dk, ib);
    if (e == hipSuccess) e = hipMemcpy(dz, z, fb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dt, t, fb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dtk, tk, ib, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(devtail_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, dz, dt, dtk, n, dm, dk, dy);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(m, dm, fb, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(k, dk, ib, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(y, dy, fb, hipMemcpyDeviceToHost);
    (void)hipFree(dz);
    (void)hipFree(dt);
    (void)hipFree(dm);
    (void)hipFree(dy);
    (void)hipFree(dtk);
    (void)hipFree(dk);
    return (int)e;
}
