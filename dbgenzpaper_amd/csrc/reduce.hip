// FP64 reductions of the integration: the per-shift sum of the tile partials
// with compensated (Neumaier) summation, the running-mean merge of
// src/qmc.jl:434-442 and the mean / 3*sd over the shifts of
// src/qmc.jl:303-312, 814-815.  Every value is carried as a mantissa and a
// binary exponent, so probabilities far below the Float64 range keep their
// log value.
#include <float.h>

#include "dbg_device.h"

namespace dbg {

constexpr int kRedThreads = 256;

__device__ __forceinline__ void neumaier_add(double& sum, double& comp, double v)
{
    const double t = sum + v;
    if (fabs(sum) >= fabs(v))
        comp += (sum - t) + v;
    else
        comp += (v - t) + sum;
    sum = t;
}

// One block per shift.
__global__ __launch_bounds__(kRedThreads) void reduce_reps_kernel(
    const double* __restrict__ part_m, const int* __restrict__ part_e, int64_t tpr,
    int64_t n_pts0, int64_t m, int sparseqmc_merge, double* rep_m, int* rep_e)
{
    __shared__ int s_e[kRedThreads / 64];
    __shared__ double s_s[kRedThreads], s_c[kRedThreads];
    const int k = blockIdx.x;
    const int tid = threadIdx.x;
    const double* pm = part_m + (size_t)k * tpr;
    const int* pe = part_e + (size_t)k * tpr;

    int e = INT32_MIN;
    for (int64_t t = tid; t < tpr; t += kRedThreads)
        if (pm[t] != 0.0) e = max(e, pe[t]);
    e = wave_max_int(e);
    if ((tid & 63) == 0) s_e[tid >> 6] = e;
    __syncthreads();
    int E = s_e[0];
    for (int w = 1; w < kRedThreads / 64; ++w) E = max(E, s_e[w]);
    if (E == INT32_MIN) E = 0;

    double sum = 0.0, comp = 0.0;
    for (int64_t t = tid; t < tpr; t += kRedThreads)
        neumaier_add(sum, comp, pm[t] != 0.0 ? ldexp(pm[t], pe[t] - E) : pm[t]);
    s_s[tid] = sum;
    s_c[tid] = comp;
    __syncthreads();
    if (tid == 0) {
        double S = 0.0, Cc = 0.0;
        for (int t = 0; t < kRedThreads; ++t) {
            neumaier_add(S, Cc, s_s[t]);
            Cc += s_c[t];
        }
        S += Cc;
        // running-mean merge with the previous rounds of this shift
        const int64_t n_tot = n_pts0 + m;
        double rm = rep_m[k];
        int re = rep_e[k];
        double mean_m;
        if (!sparseqmc_merge) {
            const double inv_tot = 1.0 / (double)n_tot;
            const double old_w = (double)n_pts0 * inv_tot;
            mean_m = S * inv_tot;
            const int Ec = (rm == 0.0) ? E : (mean_m == 0.0 ? re : max(re, E));
            rm = fma(ldexp(rm, re - Ec), old_w, ldexp(mean_m, E - Ec));
            re = Ec;
        } else {
            // SparseQMC: reps * (n0 / n_tot) + sum / n_tot   (src/qmc_sparse.jl:257-260)
            mean_m = S / (double)n_tot;
            const int Ec = (rm == 0.0) ? E : (mean_m == 0.0 ? re : max(re, E));
            rm = ldexp(rm, re - Ec) * ((double)n_pts0 / (double)n_tot) + ldexp(mean_m, E - Ec);
            re = Ec;
        }
        int en = 0;
        const double mn = frexp(rm, &en);
        rep_m[k] = mn;
        rep_e[k] = (mn == 0.0 || !isfinite(mn)) ? 0 : re + en;
        if (!isfinite(mn)) rep_m[k] = rm;
    }
}

// out[0..6] = value, log_value, err3, rel_err3, mean mantissa, err mantissa, exponent
__global__ void summary_kernel(const double* rep_m, const int* rep_e, int n_reps, double* out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int E = INT32_MIN;
    for (int k = 0; k < n_reps; ++k)
        if (rep_m[k] != 0.0 && rep_e[k] > E) E = rep_e[k];
    if (E == INT32_MIN) E = 0;
    double sum = 0.0;
    for (int k = 0; k < n_reps; ++k) sum += ldexp(rep_m[k], rep_e[k] - E);
    const double result = sum / n_reps;
    double acc = 0.0;
    for (int k = 0; k < n_reps; ++k) {
        const double v = ldexp(rep_m[k], rep_e[k] - E) - result;
        acc += v * v / n_reps;
    }
    const double err = 3.0 * sqrt(acc / (n_reps - 1));
    out[0] = ldexp(result, E);
    out[1] = result > 0 ? log(result) + E * 0.69314718055994530942 : (result == 0 ? -INFINITY : NAN);
    out[2] = ldexp(err, E);
    out[3] = err / result;
    out[4] = result;
    out[5] = err;
    out[6] = (double)E;
}

hipError_t launch_reduce_reps(const double* part_m, const int* part_e, int64_t tiles_per_rep,
                              int n_reps, int64_t n_pts0, int64_t m, int sparseqmc_merge,
                              double* rep_m, int* rep_e, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(reduce_reps_kernel, dim3(n_reps), dim3(kRedThreads), 0, s, part_m, part_e,
                       tiles_per_rep, n_pts0, m, sparseqmc_merge, rep_m, rep_e);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(summary_kernel, dim3(1), dim3(64), 0, s, rep_m, rep_e, n_reps, out);
    return hipGetLastError();
}

}  // namespace dbg
