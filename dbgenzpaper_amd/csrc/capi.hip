// C ABI of libdbgenz.so: plan objects, the factorization drivers and the
// qmc_pnorm! driver (src/qmc.jl:783-850) on top of the kernels.
#include <float.h>
#include <math.h>
#include <string.h>

#include <string>
#include <algorithm>
#include <numeric>
#include <vector>
#include <mutex>

#include <rocblas/rocblas.h>

#include "../../include/dbgenz.h"
#include "dbg_internal.h"

using namespace dbg;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess)                                                                  \
            return fail(DBG_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_));       \
    } while (0)

template <typename P>
hipError_t dalloc(P** p, size_t bytes)
{
    *p = nullptr;
    return hipMalloc((void**)p, bytes > 0 ? bytes : 16);
}

// Window boundaries of the reference's loop nest (iv -> iii -> ii blocks),
// src/cholesky.jl:482-491.
std::vector<std::pair<int, int>> pivot_windows(int n, int bs2, int bs3, int bs4)
{
    std::vector<std::pair<int, int>> w;
    for (int iv0 = 0; iv0 < n; iv0 += bs4) {
        const int iv1 = std::min(n, iv0 + bs4);
        for (int iii0 = iv0; iii0 < iv1; iii0 += bs3) {
            const int iii1 = std::min(iv1, iii0 + bs3);
            for (int ii0 = iii0; ii0 < iii1; ii0 += bs2) w.push_back({ii0, std::min(iii1, ii0 + bs2)});
        }
    }
    return w;
}

// Lattice generating vector q_g = frac(sqrt(p_g)), p_1 = 2, p_2 = 3, ...
// (get_sqrt_primes!, src/qmc_generators.jl:300-310).  Setup-time host work,
// like the reference's constructor; at least one entry.
std::vector<double> sqrt_prime_fracs(int nd)
{
    std::vector<double> qh(std::max(nd, 1), 0.0);
    int found = 0;
    long p = 1;
    while (found < nd) {
        long c = p + 1;
        for (;; ++c) {
            bool is_p = c >= 2;
            for (long f = 2; f * f <= c && is_p; ++f)
                if (c % f == 0) is_p = false;
            if (is_p) break;
        }
        p = c;
        const double sq = sqrt((double)p);
        qh[found++] = sq - floor(sq);
    }
    return qh;
}

// q_g in 64-bit fixed point, floor(q_g 2^64) (q_g in (0, 1))
std::vector<uint64_t> fixed_q64(const std::vector<double>& q)
{
    std::vector<uint64_t> r(q.size());
    for (size_t i = 0; i < q.size(); ++i) r[i] = (uint64_t)ldexp(q[i] - floor(q[i]), 64);
    return r;
}

// shifts in 32-bit fixed point, floor(frac(shift) 2^32)
std::vector<uint32_t> fixed_shift32(const double* sh, size_t cnt)
{
    std::vector<uint32_t> r(cnt);
    for (size_t i = 0; i < cnt; ++i) r[i] = (uint32_t)(uint64_t)ldexp(sh[i] - floor(sh[i]), 32);
    return r;
}

struct CholBuffers {
    double* M = nullptr;  // n x n, lda = n
    double* a = nullptr;
    double* b = nullptr;
    double* scale = nullptr;
    double* s_vec = nullptr;
    double* eps0 = nullptr;
    int32_t* perm = nullptr;
    int32_t* rank = nullptr;

    int alloc(int n)
    {
        const size_t nn = (size_t)n * n;
        HIPCHK(dalloc(&M, nn * sizeof(double)));
        HIPCHK(dalloc(&a, n * sizeof(double)));
        HIPCHK(dalloc(&b, n * sizeof(double)));
        HIPCHK(dalloc(&scale, n * sizeof(double)));
        HIPCHK(dalloc(&s_vec, n * sizeof(double)));
        HIPCHK(dalloc(&eps0, sizeof(double)));
        HIPCHK(dalloc(&perm, n * sizeof(int32_t)));
        HIPCHK(dalloc(&rank, sizeof(int32_t)));
        return DBG_OK;
    }
    void release()
    {
        (void)hipFree(M); (void)hipFree(a); (void)hipFree(b); (void)hipFree(scale);
        (void)hipFree(s_vec); (void)hipFree(eps0); (void)hipFree(perm); (void)hipFree(rank);
        M = a = b = scale = s_vec = eps0 = nullptr;
        perm = rank = nullptr;
    }
};

int check_block_sizes(int bs1, int bs2, int bs3, int bs4)
{
    if (bs1 < 1 || bs2 < 1 || bs3 < 1 || bs4 < 1)
        return fail(DBG_ERR_ARG, "Cholesky block sizes must be positive");
    return DBG_OK;
}

// Copy the inputs (host or device memory, hipMemcpyDefault) and factor.
// pivot: 0 cholesky_classic!, 1 cholesky_genz!, 2 cholesky_rowmax!.
// rowmax_tol: the rank check of the row-max pivoting is m_ii > rowmax_tol
// (0 for cholesky_rowmax! as the reference has it).  keep_stop: leave the
// solves after a rank stop undone, as the reference does, which reads the
// rank back after every pivot window; a caller that discards the factor of a
// rank-deficient matrix passes false and the windows follow each other
// without a host round trip (the later pivot blocks then exit at once).
int factor_on_device(CholBuffers& B, int n, const double* M_in, int64_t lda, const double* a_in,
                     const double* b_in, int bs1, int bs2, int bs3, int bs4, int pivot,
                     hipStream_t s, int32_t* rank_out, double rowmax_tol = 0.0, bool keep_stop = true)
{
    int rc = check_block_sizes(bs1, bs2, bs3, bs4);
    if (rc) return rc;
    HIPCHK(hipMemcpy2DAsync(B.M, (size_t)n * sizeof(double), M_in, (size_t)lda * sizeof(double),
                            (size_t)n * sizeof(double), n, hipMemcpyDefault, s));
    HIPCHK(hipMemcpyAsync(B.a, a_in, n * sizeof(double), hipMemcpyDefault, s));
    HIPCHK(hipMemcpyAsync(B.b, b_in, n * sizeof(double), hipMemcpyDefault, s));
    std::vector<int32_t> iota(n);
    for (int i = 0; i < n; ++i) iota[i] = i;
    HIPCHK(hipMemcpyAsync(B.perm, iota.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemsetAsync(B.s_vec, 0, n * sizeof(double), s));
    CholArgs C{B.M, (int64_t)n, n, B.a, B.b, B.scale, B.perm, B.s_vec, B.eps0, B.rank};
    HIPCHK(chol_standardize(C, false, s));
    if (pivot == 2) {
        // the window kernels compare the row-max pivot with the eps0 slot
        HIPCHK(hipMemcpyAsync(B.eps0, &rowmax_tol, sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    const int nb = std::max(1, std::min(bs1, 64));
    for (auto& w : pivot_windows(n, bs2, bs3, bs4))
        if (w.second - w.first > chol_max_window())
            return fail(DBG_ERR_ARG, "Cholesky pivot window (min(chol_block_size2, 1024-aligned blocks)) above " +
                                         std::to_string(chol_max_window()));
    // the reference's loop nest iv (bs4) > iii (bs3) > ii (bs2): each window is
    // factored, then solved and applied up to the end of its iii block; a
    // finished iii block is solved and applied up to the end of its iv block,
    // a finished iv block up to n (src/cholesky.jl:482-491,589-624).  A pivoted
    // factorization that stops for rank leaves all later solves undone, as the
    // reference's break out of the loop nest does: the rank is read back after
    // each window.
    int32_t rank = n;
    bool stopped = false;
    for (int iv0 = 0; iv0 < n && !stopped; iv0 += bs4) {
        const int iv1 = std::min(n, iv0 + bs4);
        for (int iii0 = iv0; iii0 < iv1 && !stopped; iii0 += bs3) {
            const int iii1 = std::min(iv1, iii0 + bs3);
            for (int ii0 = iii0; ii0 < iii1; ii0 += bs2) {
                const int ii1 = std::min(iii1, ii0 + bs2);
                HIPCHK(chol_window(C, ii0, ii1, nb, pivot, s));
                if (pivot && keep_stop) {
                    HIPCHK(hipMemcpyAsync(&rank, B.rank, sizeof(int32_t), hipMemcpyDeviceToHost, s));
                    HIPCHK(hipStreamSynchronize(s));
                    if (rank < n) {
                        stopped = true;
                        break;
                    }
                }
                HIPCHK(chol_trsm_syrk(C, ii0, ii1, ii1, iii1, s));
            }
            if (!stopped) HIPCHK(chol_trsm_syrk(C, iii0, iii1, iii1, iv1, s));
        }
        if (!stopped) HIPCHK(chol_trsm_syrk(C, iv0, iv1, iv1, n, s));
    }
    HIPCHK(chol_finish(C, s));
    HIPCHK(hipMemcpyAsync(&rank, B.rank, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *rank_out = rank;
    return DBG_OK;
}

int chol_host(double* M, int64_t lda, int32_t n, double* a, double* b, int32_t bs1, int32_t bs2,
              int32_t bs3, int32_t bs4, int32_t* perm, double* scale, int32_t* rank, int pivot)
{
    if (n < 1 || lda < n || !M || !a || !b) return fail(DBG_ERR_ARG, "invalid matrix arguments");
    hipStream_t s;
    HIPCHK(hipStreamCreate(&s));
    CholBuffers B;
    int rc = B.alloc(n);
    int32_t rk = n;
    if (!rc) rc = factor_on_device(B, n, M, lda, a, b, bs1, bs2, bs3, bs4, pivot, s, &rk);
    if (!rc) {
        hipError_t e = hipMemcpy2D(M, (size_t)lda * sizeof(double), B.M, (size_t)n * sizeof(double),
                                   (size_t)n * sizeof(double), n, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(a, B.a, n * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(b, B.b, n * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess && scale) e = hipMemcpy(scale, B.scale, n * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess && perm) e = hipMemcpy(perm, B.perm, n * sizeof(int32_t), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(DBG_ERR_DEVICE, std::string("copy back: ") + hipGetErrorString(e));
        if (rank) *rank = rk;
    }
    B.release();
    (void)hipStreamDestroy(s);
    return rc;
}

}  // namespace

// ---------------------------------------------------------------------------
// Fixed scratch of the allocation-free factorization: the conditional means
// s_vec of one pivot window and eps0 (kNoAllocMaxN + 1 doubles, 4 KB), one
// copy per device for the life of the process.  Host code never passes the
// symbol itself to a kernel: it takes the device address from
// hipGetSymbolAddress once per device (noalloc_scratch below).
// ---------------------------------------------------------------------------
constexpr int kNoAllocMaxN = 511;
static __device__ __align__(256) double scratch[kNoAllocMaxN + 1];

__global__ void noalloc_init_kernel(double* s_vec, int n, int32_t* perm)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        s_vec[i] = 0.0;
        perm[i] = i;
    }
}

namespace {

constexpr int kNoAllocMaxDev = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kNoAllocMaxDev] = {};

// device address of `scratch` on the current device (looked up on first use)
int noalloc_scratch(double** out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDev) return fail(DBG_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return DBG_OK;
}

}  // namespace

struct dbg_plan {
    dbg_opts o{};
    int n = 0, nd = 0;
    bool mixed = false;
    size_t tsz = 8;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double* sig = nullptr;   // device copies of the inputs Sigma (lda = n), a, b
    double* a0 = nullptr;
    double* b0 = nullptr;
    CholBuffers F;
    int32_t rank = 0;
    // rank-deficient dense plans (qmc_lowrank.hip): F holds the reordered
    // factor of adapt_low_rank!, these its level structure
    struct LowRank {
        bool on = false;
        int r = 0, k0 = 0, grid = 0;
        std::vector<int32_t> kjs;     // last column of each level (0-based)
        int32_t* beg = nullptr;       // r + 1 block starts
        double* a = nullptr;          // bounds of the columns (standardized, scaled)
        double* b = nullptr;
        double* inv = nullptr;        // 1 / (U[i, c] sqrt 2)
        int32_t* end = nullptr;       // adapt_low_rank! outputs and scratch
        int32_t* k0d = nullptr;
        double* tmp = nullptr;
        double* Y = nullptr;          // grid slots of (r - 1) x tile values
        int64_t slot = 0, tpr = 0, tiles = 0, y_cap = 0;
        std::vector<void*> all() const { return {beg, a, b, inv, end, k0d, tmp, Y}; }
    } L;
    // packed layouts
    void* Uk = nullptr;
    float* ucol32 = nullptr;  // FP32 copy of the factor for the mixed walk
    void* uc16 = nullptr;     // split FP16 factor of the FP16 matrix-core contraction
    bool split = false;       // mixed dense path on the split-FP16 kernel
    int64_t ldu = 0;
    void* ak = nullptr;
    void* bk = nullptr;
    void* invk = nullptr;
    double* first = nullptr;
    double first_h[6] = {0, 0, 0, 0, 0, 0};
    int64_t* colcnt = nullptr;
    int64_t* colptr = nullptr;
    int32_t* rowval = nullptr;
    void* nzval = nullptr;
    int64_t nnz = 0;
    int* reach_d = nullptr;  // pattern reach (largest g - r read by column g + 1)
    int ring = 0;            // LDS ring rows of the sparse kernel, 0: global workspace
    int ring_grid = 0;       // resident blocks of that ring kernel
    bool a_inf = false;      // every lower bound is -inf
    // points
    double* q = nullptr;
    double* shifts_f = nullptr;
    uint64_t* q64 = nullptr;     // mixed path: q in 64-bit fixed point
    uint32_t* shift32 = nullptr; // mixed path: Richtmyer shifts in 32-bit fixed point
    uint64_t* shifts_u = nullptr;
    uint32_t* V = nullptr;
    bool shifts_set = false, dirs_set = false;
    int64_t state = 0;
    uint64_t sobol_skip = 0;
    // integration workspace
    void* Y = nullptr;
    int64_t y_slot = 0;
    int grid = 0;
    int ncu = 0;
    int tile = 256;
    bool use_mfma = false;
    int64_t tiles_per_rep = 0, n_tiles = 0;
    double* part_m = nullptr;
    int* part_e = nullptr;
    double* rep_m = nullptr;
    int* rep_e = nullptr;
    double* out = nullptr;
    // FP64 re-evaluation of the heaviest tiles (mixed dense plans): K tiles per
    // shift, the FP64 kernel's layouts of the FP64-truncated factor, its tiles
    // (ratio per mixed tile) and partials, the picks and the unpicked shares
    struct Refine {
        int K = 0;
        int ratio = 1;
        bool mfma = false;
        int grid = 0;
        int64_t ldu = 0, tpr = 0, n_tiles = 0, y_slot = 0;
        double* ucol = nullptr;   // n x n, column-major (the FP64 factor)
        double* uk = nullptr;     // packed rows, ldu
        double* ak = nullptr;
        double* bk = nullptr;
        double* invk = nullptr;
        double* first = nullptr;
        double* Y = nullptr;      // grid slots of y_slot
        double* part_m = nullptr; // n_tiles
        int* part_e = nullptr;
        int64_t* sel = nullptr;   // n_reps x K
        int64_t* list = nullptr;  // n_reps x K x ratio
        double* frac = nullptr;   // n_reps
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        std::vector<void*> all() const { return {ucol, uk, ak, bk, invk, first, Y, part_m, part_e, sel, list, frac}; }
    } R;
    // adjoint buffers (allocated on the first dbg_qmc_grad)
    struct Grad {
        bool ready = false;
        int64_t budget = kGradRecordBudget;  // bytes of y and s records
        int64_t bytes = 0;      // device memory held by the buffers below
        int tile = 0, grid = 0;
        int64_t tiles_per_rep = 0, n_tiles = 0;
        // the samples run in chunks of chunk_tiles tiles (the last may be
        // shorter), each recorded, swept back and folded into the sums
        int64_t chunk_tiles = 0, rs = 0;  // rs: record row stride, chunk_tiles * tile + kRecPad
        int n_schunks = 0;
        void* yrec = nullptr;   // nd x rs (T): y of every sample of a chunk
        void* srec = nullptr;   // nd x rs (T): s_{g+1}, then sbar_{g+1}
        void* wm = nullptr;     // rs (T)
        int* we = nullptr;      // rs
        void* ywork = nullptr;  // grid x nd x tile (T): reverse workspace
        double* part_m = nullptr;  // chunk_tiles
        int* part_e = nullptr;
        double* norm = nullptr; // E, 1/S, log P, S of a chunk (log P of all at the end)
        double* run = nullptr;  // E, S of the chunks folded so far; alpha, beta
        double* acc = nullptr;  // 3 n: d log P / d a', b', U_rr (all chunks)
        double* cacc = nullptr; // 3 n: the same for one chunk
        void* ufk = nullptr;    // flipped factor, packed rows (T)
        void* ufc = nullptr;    // flipped factor, column-major (T)
        void* uc16f = nullptr;  // flipped factor, split FP16 layout (mixed split path)
        void* part = nullptr;   // split-K partial tiles of Y^T Sbar (T)
        int2* pairs = nullptr;  // upper-triangle tile pairs
        int n_pairs = 0, n_chunks = 0;
        int64_t kchunk = 0;
        double* ub = nullptr;   // n x n: d log P / d U
        double* x = nullptr;    // n x n work
        double* g = nullptr;    // n x n: d log P / d Sigma
        double* gab = nullptr;  // 2 n: d log P / d a, b
        std::vector<void*> all() const
        {
            return {yrec, srec, wm, we, ywork, part_m, part_e, norm, run, acc, cacc, ufk, ufc, uc16f, part, pairs,
                    ub, x, g, gab};
        }
        void free_all()
        {
            for (void* p : all()) (void)hipFree(p);
            const int64_t keep = budget;
            *this = Grad{};
            budget = keep;
        }
    } G;

    void release()
    {
        G.free_all();
        for (void* p : R.all()) (void)hipFree(p);
        for (void* p : L.all()) (void)hipFree(p);
        if (R.ev0) (void)hipEventDestroy(R.ev0);
        if (R.ev1) (void)hipEventDestroy(R.ev1);
        F.release();
        void* ptrs[] = {sig, a0, b0, Uk, ucol32, uc16, ak, bk, invk, first, colcnt, colptr, rowval, nzval, reach_d, q, shifts_f,
                        q64, shift32, shifts_u, V, Y, part_m, part_e, rep_m, rep_e, out};
        for (void* p : ptrs) (void)hipFree(p);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

#define BLASCHK(expr)                                                                          \
    do {                                                                                       \
        rocblas_status st_ = (expr);                                                           \
        if (st_ != rocblas_status_success)                                                     \
            return fail(DBG_ERR_DEVICE, std::string("rocBLAS: ") + rocblas_status_to_string(st_)); \
    } while (0)

// The route of a dense plan whose Genz factorization stops short of full
// rank (no reference counterpart: QMCData throws, src/qmc.jl:195-197):
// cholesky_rowmax! with the rank check m_ii > n eps (LAPACK dpstrf's default
// tolerance), rows 0 .. r-1 of the dependent columns completed from Sigma
// (U_r^T X = C_p[0:r, r:n]), the standardized bounds permuted, and
// adapt_low_rank! with eps1 = sqrt(eps): its blocks of columns, one per
// latent variable, are integrated in FP64 by the level loop of
// qmc_lowrank.hip.  A factor of full rank under the row-max pivoting runs the
// same loop with one column per block.
int lowrank_factor(dbg_plan* P)
{
    const dbg_opts& o = P->o;
    const int n = P->n;
    hipStream_t s = P->stream;
    auto& L = P->L;
    int32_t r = n;
    int rc = factor_on_device(P->F, n, P->sig, n, P->a0, P->b0, o.chol_block_size, o.chol_block_size2,
                              o.chol_block_size3, o.chol_block_size4, 2, s, &r, n * DBL_EPSILON);
    if (rc) return rc;
    if (r < 1) return fail(DBG_ERR_RANK, "the covariance matrix has rank 0");
    if (L.beg == nullptr) {
        HIPCHK(dalloc(&L.beg, (size_t)(n + 1) * sizeof(int32_t)));
        HIPCHK(dalloc(&L.a, (size_t)n * sizeof(double)));
        HIPCHK(dalloc(&L.b, (size_t)n * sizeof(double)));
        HIPCHK(dalloc(&L.inv, (size_t)n * sizeof(double)));
        HIPCHK(dalloc(&L.end, (size_t)n * sizeof(int32_t)));
        HIPCHK(dalloc(&L.k0d, sizeof(int32_t)));
        HIPCHK(dalloc(&L.tmp, (size_t)n * n * sizeof(double)));
    }
    // rows 0 .. r-1 of the dependent columns: C_p[0:r, r:n], then U_r^{-T} of it
    HIPCHK(lowrank_fill(P->F.M, n, r, P->sig, P->F.perm, P->F.scale, s));
    if (r < n) {
        rocblas_handle h = (rocblas_handle)blas_handle_void(s);
        if (!h) return fail(DBG_ERR_DEVICE, "rocBLAS handle");
        const double one = 1.0;
        BLASCHK(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                              rocblas_diagonal_non_unit, r, n - r, &one, P->F.M, n, P->F.M + (size_t)r * n, n));
    }
    HIPCHK(lowrank_bounds(P->F.a, P->F.b, P->F.perm, n, L.a, L.b, s));
    HIPCHK(adapt_low_rank(P->F.M, n, n, L.a, L.b, P->F.perm, r, std::sqrt(DBL_EPSILON), L.end, L.tmp, L.k0d, s));
    L.kjs.assign(r, 0);
    int32_t k0 = 0;
    HIPCHK(hipMemcpyAsync(L.kjs.data(), L.end, (size_t)r * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&k0, L.k0d, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::vector<int32_t> beg(r + 1, 0);
    for (int i = 0; i < r; ++i) beg[i + 1] = L.kjs[i] + 1;
    for (int i = 0; i < r; ++i)
        if (beg[i + 1] <= beg[i]) return fail(DBG_ERR_DEVICE, "adapt_low_rank: empty level block");
    if (beg[r] != n - k0) return fail(DBG_ERR_DEVICE, "adapt_low_rank: level blocks do not cover the columns");
    HIPCHK(hipMemcpyAsync(L.beg, beg.data(), (size_t)(r + 1) * sizeof(int32_t), hipMemcpyHostToDevice, s));
    HIPCHK(lowrank_prep(P->F.M, n, L.beg, r, L.a, L.b, L.inv, P->first, o.phi_form, s));
    HIPCHK(hipMemcpyAsync(P->first_h, P->first, 6 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // integration layout: tiles of lowrank_tile() points, (r - 1) y per point
    const int tile = lowrank_tile();
    L.tpr = (o.m + tile - 1) / tile;
    L.tiles = L.tpr * o.n_reps;
    // (at a refactor: the plan's partial sums were sized by its dense tiles, never longer)
    if (P->n_tiles > 0 && L.tiles > P->n_tiles)
        return fail(DBG_ERR_DEVICE, "rank-deficient route: partial sums do not fit");
    L.grid = (int)std::min<int64_t>(L.tiles, (int64_t)P->ncu * 4);
    L.slot = (int64_t)std::max(r - 1, 1) * tile;
    const int64_t need = (int64_t)L.grid * L.slot;
    if (need > L.y_cap) {
        (void)hipFree(L.Y);
        L.Y = nullptr;
        HIPCHK(dalloc(&L.Y, (size_t)need * sizeof(double)));
        L.y_cap = need;
    }
    L.r = r;
    L.k0 = k0;
    L.on = true;
    P->rank = r;
    return DBG_OK;
}

// Factor the device-resident inputs and rebuild the kernel layouts (the
// QMCData constructor's work, src/qmc.jl:182-199).
int plan_factor(dbg_plan* P)
{
    const dbg_opts& o = P->o;
    const int n = P->n;
    hipStream_t s = P->stream;
    const bool pivot = (o.kind == DBG_KIND_DENSE);
    P->L.on = false;
    // a rank stop sends the plan to lowrank_factor, which factors Sigma again:
    // this factor is then discarded, so no host round trip per window
    int rc = factor_on_device(P->F, n, P->sig, n, P->a0, P->b0, o.chol_block_size, o.chol_block_size2,
                              o.chol_block_size3, o.chol_block_size4, pivot ? 1 : 0, s, &P->rank, 0.0, false);
    if (rc) return rc;
    if (pivot && P->rank != n) return lowrank_factor(P);
    // truncate_matrixU!: |U_ij| < eps(T) / (9n) -> 0.  A refining mixed plan
    // first keeps the factor truncated as in Float64 for its FP64 tiles.
    if (P->R.K > 0) {
        HIPCHK(truncate_upper(P->F.M, n, n, DBL_EPSILON / (double)(n * 9), s));
        HIPCHK(hipMemcpyAsync(P->R.ucol, P->F.M, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, s));
        HIPCHK(pack_dense(P->R.ucol, n, n, P->F.a, P->F.b, false, P->R.uk, P->R.ldu, P->R.ak, P->R.bk, P->R.invk,
                          P->R.first, DBG_PHI_ERF, s));
    }
    const double min_val = P->mixed ? (double)(FLT_EPSILON / (float)(n * 9)) : DBL_EPSILON / (double)(n * 9);
    HIPCHK(truncate_upper(P->F.M, n, n, min_val, s));
    const size_t T = P->tsz;
    if (o.kind != DBG_KIND_DENSE) {
        HIPCHK(csc_count(P->F.M, n, n, P->colcnt, s));
        HIPCHK(exclusive_scan_i64(P->colcnt, P->colptr, n, s));
        int64_t nnz = 0;
        HIPCHK(hipMemcpyAsync(&nnz, P->colptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (nnz > P->nnz || P->rowval == nullptr) {
            (void)hipFree(P->rowval);
            (void)hipFree(P->nzval);
            P->rowval = nullptr;
            P->nzval = nullptr;
            HIPCHK(dalloc(&P->rowval, (size_t)std::max<int64_t>(nnz, 1) * sizeof(int32_t)));
            HIPCHK(dalloc(&P->nzval, (size_t)std::max<int64_t>(nnz, 1) * T));
        }
        P->nnz = nnz;
        HIPCHK(csc_fill(P->F.M, n, n, P->colptr, P->mixed, P->rowval, P->nzval, s));
        int reach = -1;
        HIPCHK(csc_reach(P->colptr, P->rowval, n, P->reach_d, s));
        HIPCHK(hipMemcpyAsync(&reach, P->reach_d, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        P->ring = sparse_ring(P->mixed, reach);
        const int gen = (o.kind == DBG_KIND_SPARSEQMC) ? kGenSobol : o.generator;
        const int form = (o.phi_form == DBG_PHI_TAIL) ? kPhiTail : kPhiErf;
        if (P->ring > 0) HIPCHK(sparse_grid(P->mixed, gen, form, P->ring, P->ncu, &P->ring_grid));
    }
    if (o.kind == DBG_KIND_DENSE && P->mixed) {
        if (P->ucol32 == nullptr) HIPCHK(dalloc(&P->ucol32, (size_t)n * n * sizeof(float)));
        HIPCHK(convert_f32(P->F.M, (size_t)n * n, P->ucol32, s));
        if (P->split) HIPCHK(pack_split16(P->ucol32, n, P->uc16, s));
    }
    HIPCHK(pack_dense(P->F.M, n, n, P->F.a, P->F.b, P->mixed, P->Uk, P->ldu, P->ak, P->bk, P->invk,
                      P->first, o.phi_form == DBG_PHI_ERFC ? DBG_PHI_ERFC : DBG_PHI_ERF, s));
    HIPCHK(hipMemcpyAsync(P->first_h, P->first, 6 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return DBG_OK;
}

int plan_build(dbg_plan* P, const double* Sigma, int64_t lda, const double* a, const double* b)
{
    const dbg_opts& o = P->o;
    const int n = P->n;
    hipStream_t s = P->stream;
    int rc = P->F.alloc(n);
    if (rc) return rc;
    HIPCHK(dalloc(&P->sig, (size_t)n * n * sizeof(double)));
    HIPCHK(dalloc(&P->a0, n * sizeof(double)));
    HIPCHK(dalloc(&P->b0, n * sizeof(double)));
    // Sigma in host or device memory (a covariance built by dbg_cov_build stays on the device)
    HIPCHK(hipMemcpy2DAsync(P->sig, (size_t)n * sizeof(double), Sigma, (size_t)lda * sizeof(double),
                            (size_t)n * sizeof(double), n, hipMemcpyDefault, s));
    HIPCHK(hipMemcpyAsync(P->a0, a, n * sizeof(double), hipMemcpyHostToDevice, s));
    P->a_inf = true;
    for (int i = 0; i < n; ++i) P->a_inf = P->a_inf && std::isinf(a[i]) && a[i] < 0;
    HIPCHK(hipMemcpyAsync(P->b0, b, n * sizeof(double), hipMemcpyHostToDevice, s));

    const int nd = P->nd;
    const size_t T = P->tsz;
    HIPCHK(dalloc(&P->ak, n * T));
    HIPCHK(dalloc(&P->bk, n * T));
    HIPCHK(dalloc(&P->invk, n * T));
    HIPCHK(dalloc(&P->first, 8 * sizeof(double)));
    if (o.kind == DBG_KIND_DENSE) {
        // rows padded so a panel of targets past n reads zeros
        P->ldu = (int64_t)n + dense_mfma_panel(P->mixed) + 32;
        HIPCHK(dalloc(&P->Uk, (size_t)std::max(nd, 1) * P->ldu * T));
    } else {
        HIPCHK(dalloc(&P->colcnt, (size_t)n * sizeof(int64_t)));
        HIPCHK(dalloc(&P->colptr, (size_t)(n + 1) * sizeof(int64_t)));
        HIPCHK(dalloc(&P->reach_d, sizeof(int)));
    }
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    P->ncu = ncu;
    P->use_mfma = (o.kind == DBG_KIND_DENSE) && n >= kMfmaMinDim;
    // the FP16 matrix-core contraction: mixed dense path
    P->split = P->use_mfma && P->mixed && o.contraction != DBG_CONTRACT_FP32;
    if (P->split)
        HIPCHK(dalloc(&P->uc16, (size_t)split16_cols(n) * (size_t)split16_col_stride(n) * sizeof(uint16_t)));
    // tiles of the mixed kernel (for the refinement's tile arithmetic)
    const int tile_mixed = P->split ? dense_split_tile() : (P->use_mfma ? dense_mfma_tile(true) : dense_valu_tile());
    const int64_t tpr_mixed = (o.m + tile_mixed - 1) / tile_mixed;
    P->R.K = (P->mixed && o.kind == DBG_KIND_DENSE && nd > 0)
                 ? (int)std::min<int64_t>(o.refine_tiles < 0 ? kDefaultRefineTiles : o.refine_tiles, tpr_mixed)
                 : 0;
    if (P->R.K > 0) {
        auto& R = P->R;
        R.mfma = n >= kMfmaMinDim;
        const int t64 = R.mfma ? dense_mfma_refine_tile() : dense_valu_tile();
        if (tile_mixed % t64 != 0) return fail(DBG_ERR_ARG, "refinement: mixed tile not a multiple of the FP64 tile");
        R.ratio = tile_mixed / t64;
        R.ldu = (int64_t)n + dense_mfma_refine_panel() + 32;
        R.tpr = (o.m + t64 - 1) / t64;
        R.n_tiles = R.tpr * o.n_reps;
        const int64_t nlist = (int64_t)o.n_reps * R.K * R.ratio;
        R.grid = (int)std::min<int64_t>(nlist, (int64_t)ncu * (R.mfma ? 1 : 2));
        R.y_slot = (int64_t)nd * t64;
        HIPCHK(dalloc(&R.ucol, (size_t)n * n * sizeof(double)));
        HIPCHK(dalloc(&R.uk, (size_t)nd * R.ldu * sizeof(double)));
        HIPCHK(dalloc(&R.ak, n * sizeof(double)));
        HIPCHK(dalloc(&R.bk, n * sizeof(double)));
        HIPCHK(dalloc(&R.invk, n * sizeof(double)));
        HIPCHK(dalloc(&R.first, 8 * sizeof(double)));
        HIPCHK(dalloc(&R.Y, (size_t)R.grid * R.y_slot * sizeof(double)));
        HIPCHK(dalloc(&R.part_m, (size_t)R.n_tiles * sizeof(double)));
        HIPCHK(dalloc(&R.part_e, (size_t)R.n_tiles * sizeof(int)));
        HIPCHK(dalloc(&R.sel, (size_t)o.n_reps * R.K * sizeof(int64_t)));
        HIPCHK(dalloc(&R.list, (size_t)nlist * sizeof(int64_t)));
        HIPCHK(dalloc(&R.frac, (size_t)o.n_reps * sizeof(double)));
        HIPCHK(hipEventCreate(&R.ev0));
        HIPCHK(hipEventCreate(&R.ev1));
    }
    rc = plan_factor(P);
    if (rc) return rc;

    const std::vector<double> qh = sqrt_prime_fracs(nd);
    HIPCHK(dalloc(&P->q, qh.size() * sizeof(double)));
    HIPCHK(hipMemcpyAsync(P->q, qh.data(), qh.size() * sizeof(double), hipMemcpyHostToDevice, s));
    const std::vector<uint64_t> q64h = fixed_q64(qh);
    HIPCHK(dalloc(&P->q64, q64h.size() * sizeof(uint64_t)));
    HIPCHK(hipMemcpyAsync(P->q64, q64h.data(), q64h.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    HIPCHK(dalloc(&P->shifts_f, (size_t)std::max(nd, 1) * o.n_reps * sizeof(double)));
    HIPCHK(dalloc(&P->shift32, (size_t)std::max(nd, 1) * o.n_reps * sizeof(uint32_t)));
    HIPCHK(dalloc(&P->shifts_u, (size_t)std::max(nd, 1) * o.n_reps * sizeof(uint64_t)));
    HIPCHK(dalloc(&P->V, (size_t)std::max(nd, 1) * 32 * sizeof(uint32_t)));

    // integration workspace: one y slot per resident block
    P->tile = (o.kind != DBG_KIND_DENSE) ? sparse_tile()
                                          : (P->split ? dense_split_tile()
                                                      : (P->use_mfma ? dense_mfma_tile(P->mixed) : dense_valu_tile()));
    P->tiles_per_rep = (o.m + P->tile - 1) / P->tile;
    P->n_tiles = P->tiles_per_rep * o.n_reps;
    P->grid = (int)std::min<int64_t>(P->n_tiles, (int64_t)ncu * 2);
    P->y_slot = P->split ? dense_split_slot(nd) : (int64_t)std::max(nd, 1) * P->tile;
    HIPCHK(dalloc(&P->Y, (size_t)P->grid * P->y_slot * T));
    const int64_t n_part = std::max<int64_t>(P->n_tiles, P->L.on ? P->L.tiles : 0);
    HIPCHK(dalloc(&P->part_m, (size_t)n_part * sizeof(double)));
    HIPCHK(dalloc(&P->part_e, (size_t)n_part * sizeof(int)));
    HIPCHK(dalloc(&P->rep_m, (size_t)o.n_reps * sizeof(double)));
    HIPCHK(dalloc(&P->rep_e, (size_t)o.n_reps * sizeof(int)));
    HIPCHK(dalloc(&P->out, 8 * sizeof(double)));
    HIPCHK(hipMemsetAsync(P->rep_m, 0, (size_t)o.n_reps * sizeof(double), s));
    HIPCHK(hipMemsetAsync(P->rep_e, 0, (size_t)o.n_reps * sizeof(int), s));
    HIPCHK(hipStreamSynchronize(s));
    return DBG_OK;
}

int64_t n_of(const dbg_plan* P) { return P->n; }

PointArgs point_args(const dbg_plan* P, int64_t block)
{
    PointArgs pa{};
    pa.nd = P->nd;
    pa.q = P->q;
    pa.shifts_f = P->shifts_f;
    pa.n_base = block * P->o.m;
    pa.V = P->V;
    pa.shifts_u = P->shifts_u;
    pa.sobol_first = P->sobol_skip + (uint64_t)block * (uint64_t)P->o.m;
    const int L = P->o.n_bits;
    pa.mask = (L >= 64) ? ~0ull : ((1ull << L) - 1ull);
    pa.invscale = ldexp(1.0, -L);
    pa.sobol_bits = L;
    pa.q64 = P->q64;
    pa.shift32 = P->shift32;
    return pa;
}

int run_loop(dbg_plan* P, int64_t block, int64_t n_pts0)
{
    const dbg_opts& o = P->o;
    if (P->nd > 0 && !P->shifts_set) return fail(DBG_ERR_STATE, "shifts not set");
    if (P->nd > 0 && o.generator == DBG_GEN_SOBOL && !P->dirs_set)
        return fail(DBG_ERR_STATE, "Sobol direction integers not set");
    const int gen = (o.kind == DBG_KIND_SPARSEQMC) ? kGenSobol : o.generator;
    HIPCHK(hipEventRecord(P->ev0, P->stream));
    if (P->L.on) {
        // rank-deficient route: the level loop in FP64
        const auto& L = P->L;
        DenseArgs A{};
        A.n = P->n;
        A.m = o.m;
        A.n_reps = o.n_reps;
        A.tiles_per_rep = L.tpr;
        A.n_tiles = L.tiles;
        A.Ucol = P->F.M;
        A.ldc = P->n;
        A.ak = L.a;
        A.bk = L.b;
        A.invk = L.inv;
        for (int q = 0; q < 6; ++q) A.first[q] = P->first_h[q];
        A.pts = point_args(P, block);
        A.Y = L.Y;
        A.y_slot = L.slot;
        A.part_m = P->part_m;
        A.part_e = P->part_e;
        A.lr_beg = L.beg;
        A.lr_rank = L.r;
        const int form = o.phi_form == DBG_PHI_TAIL ? kPhiTail : (o.phi_form == DBG_PHI_ERFC ? kPhiErfc : kPhiErf);
        HIPCHK(launch_lowrank(A, gen, form, L.grid, P->stream));
        HIPCHK(hipEventRecord(P->ev1, P->stream));
    } else if (o.kind == DBG_KIND_DENSE) {
        DenseArgs A{};
        A.n = P->n;
        A.m = o.m;
        A.n_reps = o.n_reps;
        A.tiles_per_rep = P->tiles_per_rep;
        A.n_tiles = P->n_tiles;
        A.Uk = P->Uk;
        A.ldu = P->ldu;
        A.Ucol = P->mixed ? (const void*)P->ucol32 : (const void*)P->F.M;
        A.ldc = n_of(P);
        A.ak = P->ak;
        A.bk = P->bk;
        A.invk = P->invk;
        for (int q = 0; q < 6; ++q) A.first[q] = P->first_h[q];
        A.pts = point_args(P, block);
        A.Y = P->Y;
        A.y_slot = P->y_slot;
        A.part_m = P->part_m;
        A.part_e = P->part_e;
        A.a_inf = P->a_inf ? 1 : 0;
        if (P->split)
            HIPCHK(launch_dense_split(A, gen, P->uc16, P->grid, P->stream));
        else if (P->use_mfma)
            HIPCHK(launch_dense_mfma(A, P->mixed, gen, o.phi_form, P->grid, P->stream));
        else
            HIPCHK(launch_dense_valu(A, P->mixed, gen, o.phi_form, P->grid, P->stream));
        HIPCHK(hipEventRecord(P->ev1, P->stream));
        if (P->R.K > 0) {
            // the heaviest tiles of each shift again, in FP64 (tail form)
            auto& R = P->R;
            HIPCHK(hipEventRecord(R.ev0, P->stream));
            HIPCHK(refine_select(P->part_m, P->part_e, P->tiles_per_rep, o.n_reps, R.K, R.ratio, R.tpr, R.sel, R.list,
                                 R.frac, P->stream));
            DenseArgs B = A;
            B.tiles_per_rep = R.tpr;
            B.n_tiles = (int64_t)o.n_reps * R.K * R.ratio;
            B.tile_list = R.list;
            B.Uk = R.uk;
            B.ldu = R.ldu;
            B.Ucol = R.ucol;
            B.ak = R.ak;
            B.bk = R.bk;
            B.invk = R.invk;
            B.Y = R.Y;
            B.y_slot = R.y_slot;
            B.part_m = R.part_m;
            B.part_e = R.part_e;
            if (R.mfma)
                HIPCHK(launch_dense_mfma_refine(B, gen, R.grid, P->stream));
            else
                HIPCHK(launch_dense_valu(B, false, gen, kPhiTail, R.grid, P->stream));
            HIPCHK(refine_merge(R.sel, o.n_reps, R.K, R.ratio, P->tiles_per_rep, R.tpr, R.part_m, R.part_e,
                                P->part_m, P->part_e, P->stream));
            HIPCHK(hipEventRecord(R.ev1, P->stream));
        }
    } else {
        SparseArgs A{};
        A.n = P->n;
        A.m = o.m;
        A.n_reps = o.n_reps;
        A.tiles_per_rep = P->tiles_per_rep;
        A.n_tiles = P->n_tiles;
        A.colptr = P->colptr;
        A.rowval = P->rowval;
        A.nzval = P->nzval;
        A.ak = P->ak;
        A.bk = P->bk;
        A.invk = P->invk;
        for (int q = 0; q < 6; ++q) A.first[q] = P->first_h[q];
        A.tent = (o.kind == DBG_KIND_SPARSEQMC) ? 0 : 1;
        A.form = (o.phi_form == DBG_PHI_TAIL) ? kPhiTail : kPhiErf;
        A.pts = point_args(P, block);
        A.Y = P->Y;
        A.y_slot = P->y_slot;
        A.part_m = P->part_m;
        A.part_e = P->part_e;
        const int grid = P->ring > 0 ? (int)std::min<int64_t>(P->n_tiles, P->ring_grid) : P->grid;
        HIPCHK(launch_sparse(A, P->mixed, gen, P->ring, grid, P->stream));
        HIPCHK(hipEventRecord(P->ev1, P->stream));
    }
    HIPCHK(launch_reduce_reps(P->part_m, P->part_e, P->L.on ? P->L.tpr : P->tiles_per_rep, o.n_reps, n_pts0, o.m,
                              o.kind == DBG_KIND_SPARSEQMC ? 1 : 0, P->rep_m, P->rep_e, P->out,
                              P->stream));
    return DBG_OK;
}

// ---------------------------------------------------------------------------
// Adjoint (no reference counterpart): forward with records, weights, reverse
// sweep, Ubar = Y^T Sbar (matrix-core kernels, split-K over the samples of a
// chunk, FP64 sums), the Cholesky adjoint C_bar = U^{-1} Phi(U Ubar^T) U^{-T}
// (rocBLAS dgemm, dtrsm) and the map to the original a, b, Sigma.
// ---------------------------------------------------------------------------
constexpr int64_t kRecPad = 256;  // samples of padding per record row (a multiple of 32)

int grad_alloc(dbg_plan* P)
{
    auto& G = P->G;
    if (G.ready) return DBG_OK;
    const int n = P->n, nd = P->nd;
    const size_t T = P->tsz;
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    G.tile = dense_mfma_tile(P->mixed);
    G.tiles_per_rep = (P->o.m + G.tile - 1) / G.tile;
    G.n_tiles = G.tiles_per_rep * P->o.n_reps;
    // chunks of equal size (whole tiles) whose y and s records fit the budget
    {
        const int64_t per_tile = 2 * (int64_t)nd * G.tile * (int64_t)T;
        const int64_t fit = std::max<int64_t>(1, G.budget / per_tile);
        G.n_schunks = (int)((G.n_tiles + fit - 1) / fit);
        G.chunk_tiles = (G.n_tiles + G.n_schunks - 1) / G.n_schunks;
    }
    // record rows (one per index, the chunk's samples in order) padded by
    // kRecPad samples: the chunk is often a power of two of samples, and a
    // row stride of 2 MB put the 2 x 256 rows that one stage of the Ubar
    // product reads on the same HBM channels (56.2 -> 51.4 ms per gradient
    // at d = 1000)
    const int64_t chunk_samples = G.chunk_tiles * G.tile;
    G.rs = chunk_samples + kRecPad;
    G.grid = (int)std::min<int64_t>(G.chunk_tiles, (int64_t)ncu * 2);
    int64_t total = 0;
    auto alloc = [&](auto** p, size_t bytes) {
        total += (int64_t)bytes;
        return dalloc(p, bytes);
    };
    const size_t rec = (size_t)nd * G.rs * T;
    HIPCHK(alloc(&G.yrec, rec));
    HIPCHK(alloc(&G.srec, rec));
    HIPCHK(alloc(&G.wm, (size_t)chunk_samples * T));
    HIPCHK(alloc(&G.we, (size_t)chunk_samples * sizeof(int)));
    if (!P->split) HIPCHK(alloc(&G.ywork, (size_t)G.grid * nd * G.tile * T));
    HIPCHK(alloc(&G.part_m, (size_t)G.chunk_tiles * sizeof(double)));
    HIPCHK(alloc(&G.part_e, (size_t)G.chunk_tiles * sizeof(int)));
    HIPCHK(alloc(&G.norm, 4 * sizeof(double)));
    HIPCHK(alloc(&G.run, 4 * sizeof(double)));
    HIPCHK(alloc(&G.acc, (size_t)3 * n * sizeof(double)));
    HIPCHK(alloc(&G.cacc, (size_t)3 * n * sizeof(double)));
    HIPCHK(alloc(&G.ufk, (size_t)nd * P->ldu * T));
    HIPCHK(alloc(&G.ufc, (size_t)n * n * T + 256 * T));
    if (P->split)
        HIPCHK(alloc(&G.uc16f, (size_t)split16_cols(n) * (size_t)split16_col_stride(n) * sizeof(uint16_t)));
    {
        // split-K layout of Ubar = Y^T Sbar over one chunk: upper-triangle
        // pairs of row tiles (128 rows, two resident workgroups per CU; 256
        // rows and one for the split form), split-K chunks in whole rounds of
        // the resident slots (pairs x chunks a multiple of the slots: no
        // short last round)
        const int tb = ubar_tile(P->split);
        const int nt = (nd + tb - 1) / tb;
        std::vector<int2> pr;
        for (int ti = 0; ti < nt; ++ti)
            for (int tj = ti; tj < nt; ++tj) pr.push_back(make_int2(ti, tj));
        G.n_pairs = (int)pr.size();
        const int slots = ubar_wgs_per_cu(P->split) * ncu;
        const int step = slots / std::gcd(G.n_pairs, slots);
        // about six workgroups per slot (three for the split form, whose
        // 256-wide partial tiles make the FP64 chunk reduction the larger cost)
        const int per_slot = P->split ? 3 : 6;
        int want = std::max(1, (per_slot * slots + G.n_pairs - 1) / G.n_pairs);
        want = (want + step - 1) / step * step;
        G.kchunk = std::max<int64_t>(32, ((chunk_samples + want - 1) / want + 31) / 32 * 32);
        G.n_chunks = (int)((chunk_samples + G.kchunk - 1) / G.kchunk);
        HIPCHK(alloc(&G.pairs, pr.size() * sizeof(int2)));
        HIPCHK(hipMemcpy(G.pairs, pr.data(), pr.size() * sizeof(int2), hipMemcpyHostToDevice));
        HIPCHK(alloc(&G.part, (size_t)G.n_chunks * G.n_pairs * tb * tb * T));
    }
    HIPCHK(alloc(&G.ub, (size_t)n * n * sizeof(double)));
    HIPCHK(alloc(&G.x, (size_t)n * n * sizeof(double)));
    HIPCHK(alloc(&G.g, (size_t)n * n * sizeof(double)));
    HIPCHK(alloc(&G.gab, (size_t)2 * n * sizeof(double)));
    G.bytes = total;
    G.ready = true;
    return DBG_OK;
}

int run_grad(dbg_plan* P)
{
    auto& G = P->G;
    const dbg_opts& o = P->o;
    const int n = P->n, nd = P->nd;
    hipStream_t s = P->stream;
    const int gen = o.generator;
    HIPCHK(hipMemsetAsync(G.acc, 0, (size_t)3 * n * sizeof(double), s));
    HIPCHK(hipMemsetAsync(G.ub, 0, (size_t)n * n * sizeof(double), s));
    HIPCHK(pack_flip(P->F.M, n, P->mixed, G.ufk, P->ldu, G.ufc, s));
    if (P->split) HIPCHK(pack_split16((const float*)G.ufc, n, G.uc16f, s));

    DenseArgs A{};
    A.n = n;
    A.m = o.m;
    A.n_reps = o.n_reps;
    A.tiles_per_rep = G.tiles_per_rep;
    A.ldu = P->ldu;
    A.ldc = n;
    A.ak = P->ak;
    A.bk = P->bk;
    A.invk = P->invk;
    for (int q = 0; q < 6; ++q) A.first[q] = P->first_h[q];
    A.pts = point_args(P, P->state);
    A.part_m = G.part_m;
    A.part_e = G.part_e;
    A.rs = G.rs;
    A.srec = G.srec;
    A.yrec = G.yrec;
    A.wm = G.wm;
    A.we = G.we;
    A.norm = G.norm;
    A.ga = G.cacc;
    A.gb = G.cacc + n;
    A.gd = G.cacc + 2 * n;
    A.a_inf = P->a_inf ? 1 : 0;
    const double total = (double)o.m * o.n_reps;
    for (int c = 0; c < G.n_schunks; ++c) {
        A.tile0 = (int64_t)c * G.chunk_tiles;
        A.n_tiles = std::min<int64_t>(G.chunk_tiles, G.n_tiles - A.tile0);
        const int grid = (int)std::min<int64_t>(A.n_tiles, G.grid);
        HIPCHK(hipMemsetAsync(G.cacc, 0, (size_t)3 * n * sizeof(double), s));
        // forward with records (tail form)
        A.mode = kModeRecord;
        A.Uk = P->Uk;
        A.Ucol = P->mixed ? (const void*)P->ucol32 : (const void*)P->F.M;
        if (P->split) {
            // the split-FP16 kernel: its own y workspace for the contraction, FP32 records
            A.Y = P->Y;
            A.y_slot = P->y_slot;  // P->Y holds P->grid slots
            HIPCHK(launch_dense_split(A, gen, P->uc16, std::min(grid, P->grid), s));
        } else {
            A.Y = G.yrec;
            A.y_slot = (int64_t)nd * G.tile;
            HIPCHK(launch_dense_adjoint(A, P->mixed, gen, grid, s));
        }
        HIPCHK(grad_norm(G.part_m, G.part_e, A.n_tiles, total, G.norm, s));
        HIPCHK(grad_fold_norm(G.norm, G.run, G.run + 2, c == 0, c + 1 == G.n_schunks, total, s));
        // reverse sweep in flipped index order
        A.mode = kModeReverse;
        A.Uk = G.ufk;
        A.Ucol = G.ufc;
        if (P->split) {
            // the split-FP16 kernel on the flipped factor
            A.Y = P->Y;
            A.y_slot = P->y_slot;
            HIPCHK(launch_dense_split(A, gen, G.uc16f, std::min(grid, P->grid), s));
        } else {
            A.Y = G.ywork;
            A.y_slot = (int64_t)nd * G.tile;
            HIPCHK(launch_dense_adjoint(A, P->mixed, gen, grid, s));
        }
        HIPCHK(grad_fold_acc(G.acc, G.cacc, 3 * n, G.run + 2, s));
        // Ubar[l, g + 1] = sum_samples y_l sbar_{g+1}: split-K on the matrix
        // cores, FP64 chunk sums, folded into the running Ubar
        HIPCHK(ubar_product(G.yrec, G.srec, G.rs, A.n_tiles * G.tile, nd, P->mixed, G.pairs, G.n_pairs, G.n_chunks,
                            G.kchunk, G.part, G.ub, n, G.run + 2, s, P->split));
    }
    rocblas_handle h = (rocblas_handle)blas_handle_void(s);
    if (!h) return fail(DBG_ERR_DEVICE, "rocBLAS handle");
    HIPCHK(sigma_chain(G.ub, n, G.acc + 2 * n, nullptr, s, 0, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
    // Cholesky adjoint: X = U Ubar^T, Phi(X), U^{-1} X, then X U^{-T}
    {
        const double one = 1.0, zero = 0.0;
        BLASCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, n, n, n, &one, P->F.M, n,
                              G.ub, n, &zero, G.x, n));
        HIPCHK(sigma_chain(nullptr, n, nullptr, G.x, s, 1, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                           nullptr, nullptr, nullptr, nullptr, nullptr));
        BLASCHK(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                              rocblas_diagonal_non_unit, n, n, &one, P->F.M, n, G.x, n));
        BLASCHK(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_transpose,
                              rocblas_diagonal_non_unit, n, n, &one, P->F.M, n, G.x, n));
    }
    HIPCHK(sigma_chain(nullptr, n, nullptr, nullptr, s, 2, G.x, P->F.perm, P->F.scale, P->sig, P->F.a, P->F.b,
                       G.acc, G.acc + n, G.g, G.gab, G.gab + n));
    return DBG_OK;
}

int read_summary(dbg_plan* P, dbg_result* r)
{
    double h[8];
    HIPCHK(hipMemcpyAsync(h, P->out, 7 * sizeof(double), hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    r->value = h[0];
    r->log_value = h[1];
    r->err3 = h[2];
    r->rel_err3 = h[3];
    return DBG_OK;
}

}  // namespace

extern "C" {

const char* dbg_last_error(void) { return g_err.c_str(); }

int32_t dbg_abi_version(void) { return 3; }

void dbg_opts_default(dbg_opts* o, int32_t precision)
{
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->chol_block_size = 1 << 6;
    o->chol_block_size2 = 1 << 9;
    o->chol_block_size3 = 1 << 10;
    o->chol_block_size4 = 1 << 11;
    o->m = 10000;
    o->block_size_i = 1 << 6;
    o->block_size_i2 = 1 << 5;
    o->block_size_j = 1 << 6;
    o->n_blocks = 1;
    o->n_reps = 12;
    o->n_bits = 32;
    o->max_pts = 10000;
    o->max_abs_err = 1e-6;
    o->precision = precision;
    o->generator = DBG_GEN_RICHTMYER;
    o->phi_form = DBG_PHI_ERF;
    o->kind = DBG_KIND_DENSE;
    o->sobol_skip = -1;
    o->refine_tiles = -1;
}

int dbg_chol_genz_f64(double* M, int64_t lda, int32_t n, double* a, double* b, int32_t bs1,
                      int32_t bs2, int32_t bs3, int32_t bs4, int32_t* perm, double* scale,
                      int32_t* rank)
{
    g_err.clear();
    return chol_host(M, lda, n, a, b, bs1, bs2, bs3, bs4, perm, scale, rank, 1);
}

int dbg_chol_rowmax_f64(double* M, int64_t lda, int32_t n, double* a, double* b, int32_t bs1,
                        int32_t bs2, int32_t bs3, int32_t bs4, int32_t* perm, double* scale,
                        int32_t* rank)
{
    g_err.clear();
    return chol_host(M, lda, n, a, b, bs1, bs2, bs3, bs4, perm, scale, rank, 2);
}

int dbg_adapt_low_rank_f64(double* M, int64_t lda, int32_t n, double* a, double* b, int32_t rank, int32_t* perm,
                           double eps1, int32_t* k_js, int32_t* k_0)
{
    g_err.clear();
    if (n < 1 || lda < n || !M || !a || !b || !perm || !k_0 || rank < 0 || rank > n || (rank > 0 && !k_js))
        return fail(DBG_ERR_ARG, "invalid adapt_low_rank arguments");
    hipStream_t s;
    HIPCHK(hipStreamCreate(&s));
    const size_t nn = (size_t)n * n;
    double *dM = nullptr, *da = nullptr, *db = nullptr, *dtmp = nullptr;
    int32_t *dperm = nullptr, *dend = nullptr, *dk0 = nullptr;
    hipError_t e = dalloc(&dM, nn * sizeof(double));
    if (e == hipSuccess) e = dalloc(&da, n * sizeof(double));
    if (e == hipSuccess) e = dalloc(&db, n * sizeof(double));
    if (e == hipSuccess) e = dalloc(&dtmp, nn * sizeof(double));
    if (e == hipSuccess) e = dalloc(&dperm, n * sizeof(int32_t));
    if (e == hipSuccess) e = dalloc(&dend, (size_t)std::max(rank, 1) * sizeof(int32_t));
    if (e == hipSuccess) e = dalloc(&dk0, sizeof(int32_t));
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(dM, (size_t)n * sizeof(double), M, (size_t)lda * sizeof(double), (size_t)n * sizeof(double),
                             n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(da, a, n * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(db, b, n * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dperm, perm, n * sizeof(int32_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = adapt_low_rank(dM, n, n, da, db, dperm, rank, eps1, dend, dtmp, dk0, s);
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(M, (size_t)lda * sizeof(double), dM, (size_t)n * sizeof(double), (size_t)n * sizeof(double),
                             n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(a, da, n * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(b, db, n * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(perm, dperm, n * sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && rank > 0) e = hipMemcpyAsync(k_js, dend, rank * sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(k_0, dk0, sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (void* p : {(void*)dM, (void*)da, (void*)db, (void*)dtmp, (void*)dperm, (void*)dend, (void*)dk0})
        (void)hipFree(p);
    (void)hipStreamDestroy(s);
    if (e != hipSuccess) return fail(DBG_ERR_DEVICE, std::string("adapt_low_rank: ") + hipGetErrorString(e));
    return DBG_OK;
}

int dbg_cov_build(int32_t kind, int32_t n, const double* params, int32_t k, const double* B, const double* psi,
                  int32_t standardize, double* Sigma, int64_t lds, int32_t sigma_on_device)
{
    g_err.clear();
    if (n < 1 || lds < n || !Sigma || kind < DBG_COV_FIXED_DENSE || kind > DBG_COV_LOWRANK)
        return fail(DBG_ERR_ARG, "invalid covariance arguments");
    if (kind != DBG_COV_LOWRANK && !params) return fail(DBG_ERR_ARG, "covariance parameters missing");
    if (kind == DBG_COV_FIXED_SPARSE && !(params[0] >= 0.0 && params[0] == std::floor(params[0])))
        return fail(DBG_ERR_ARG, "fixed_sparse: the bandwidth n_sp must be a non-negative integer");
    if (kind == DBG_COV_MATERN &&
        (!(params[0] == 0.5 || params[0] == 1.5 || params[0] == 2.5) || !(params[1] > 0.0)))
        return fail(DBG_ERR_ARG, "Matern: nu must be 0.5, 1.5 or 2.5 and dist positive");
    if (kind == DBG_COV_LOWRANK && (k < 1 || !B || !psi)) return fail(DBG_ERR_ARG, "low-rank: B (n x k) and psi needed");
    hipStream_t s;
    HIPCHK(hipStreamCreate(&s));
    double *dS = nullptr, *dB = nullptr, *dpsi = nullptr, *dscr = nullptr;
    const int64_t ld = sigma_on_device ? lds : n;
    hipError_t e = hipSuccess;
    if (sigma_on_device) dS = Sigma;
    else e = dalloc(&dS, (size_t)n * n * sizeof(double));
    if (e == hipSuccess && standardize) e = dalloc(&dscr, (size_t)n * sizeof(double));
    int rc = DBG_OK;
    if (e == hipSuccess && kind == DBG_COV_LOWRANK) {
        e = dalloc(&dB, (size_t)n * k * sizeof(double));
        if (e == hipSuccess) e = dalloc(&dpsi, (size_t)n * sizeof(double));
        if (e == hipSuccess) e = hipMemcpyAsync(dB, B, (size_t)n * k * sizeof(double), hipMemcpyDefault, s);
        if (e == hipSuccess) e = hipMemcpyAsync(dpsi, psi, (size_t)n * sizeof(double), hipMemcpyDefault, s);
        if (e == hipSuccess && cov_lowrank(n, k, dB, dpsi, dS, ld, s) != 0)
            rc = fail(DBG_ERR_DEVICE, "covariance: rocBLAS dsyrk failed");
    } else if (e == hipSuccess) {
        const int kk = kind == DBG_COV_FIXED_DENSE ? kCovFixedDense
                                                   : (kind == DBG_COV_FIXED_SPARSE ? kCovFixedSparse : kCovMatern);
        e = cov_fill(kk, n, dS, ld, params[0], kind == DBG_COV_MATERN ? params[1] : 0.0,
                     kind == DBG_COV_MATERN ? params[2] : 0.0, s);
    }
    if (e == hipSuccess && !rc && standardize) e = cov_standardize(n, dS, ld, dscr, s);
    if (e == hipSuccess && !rc && !sigma_on_device)
        e = hipMemcpy2DAsync(Sigma, (size_t)lds * sizeof(double), dS, (size_t)n * sizeof(double),
                             (size_t)n * sizeof(double), n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (!sigma_on_device) (void)hipFree(dS);
    (void)hipFree(dB);
    (void)hipFree(dpsi);
    (void)hipFree(dscr);
    (void)hipStreamDestroy(s);
    if (rc) return rc;
    if (e != hipSuccess) return fail(DBG_ERR_DEVICE, std::string("covariance: ") + hipGetErrorString(e));
    return DBG_OK;
}

int dbg_chol_genz_f64_dev_noalloc(double* M, int64_t lda, int32_t n, double* a, double* b, int32_t bs1, int32_t bs2,
                                  int32_t* perm, double* scale, int32_t* rank, void* stream)
{
    g_err.clear();
    if (!M || !a || !b || !perm || !scale || !rank || lda < n) return fail(DBG_ERR_ARG, "invalid matrix arguments");
    if (n < 1 || n > kNoAllocMaxN)
        return fail(DBG_ERR_ARG, "the allocation-free factorization takes 1 <= n <= " + std::to_string(kNoAllocMaxN));
    if (bs1 < 1 || bs2 < n)
        return fail(DBG_ERR_ARG, "the allocation-free factorization needs one pivot window (chol_block_size2 >= n)");
    double* sc = nullptr;
    int rc = noalloc_scratch(&sc);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    double* s_vec = sc;
    double* eps0 = sc + kNoAllocMaxN;
    // s_vec = 0 and perm = 0 .. n-1 by a kernel (no memset node under capture)
    hipLaunchKernelGGL(noalloc_init_kernel, dim3((n + 255) / 256), dim3(256), 0, s, s_vec, n, perm);
    HIPCHK(hipGetLastError());
    CholArgs C{M, lda, n, a, b, scale, perm, s_vec, eps0, rank};
    HIPCHK(chol_standardize(C, false, s));
    HIPCHK(chol_window(C, 0, n, std::max(1, std::min(bs1, 64)), 1, s));
    HIPCHK(chol_finish(C, s));
    return DBG_OK;
}

int dbg_chol_classic_f64(double* M, int64_t lda, int32_t n, double* a, double* b, int32_t bs1,
                         int32_t bs2, int32_t bs3, int32_t bs4, double* scale)
{
    g_err.clear();
    return chol_host(M, lda, n, a, b, bs1, bs2, bs3, bs4, nullptr, scale, nullptr, 0);
}

int dbg_plan_create(dbg_plan** plan, const dbg_opts* opts, int32_t n, const double* Sigma,
                    int64_t lda, const double* a, const double* b)
{
    g_err.clear();
    if (!plan || !opts || !Sigma || !a || !b) return fail(DBG_ERR_ARG, "null argument");
    *plan = nullptr;
    const dbg_opts& o = *opts;
    if (n < 1 || lda < n) return fail(DBG_ERR_ARG, "invalid dimension or leading dimension");
    if (o.m < 1 || o.n_reps < 1) return fail(DBG_ERR_ARG, "m and n_reps must be positive");
    if (o.precision != DBG_PREC_FP64 && o.precision != DBG_PREC_MIXED)
        return fail(DBG_ERR_ARG, "unknown precision");
    if (o.generator != DBG_GEN_RICHTMYER && o.generator != DBG_GEN_SOBOL)
        return fail(DBG_ERR_ARG, "unknown generator");
    if (o.kind < DBG_KIND_DENSE || o.kind > DBG_KIND_SPARSEQMC) return fail(DBG_ERR_ARG, "unknown kind");
    if (o.phi_form != DBG_PHI_ERF && o.phi_form != DBG_PHI_ERFC && o.phi_form != DBG_PHI_TAIL)
        return fail(DBG_ERR_ARG, "unknown phi_form");
    if (o.kind != DBG_KIND_DENSE && o.phi_form == DBG_PHI_ERFC)
        return fail(DBG_ERR_ARG, "the erfc form exists for the dense loop only");
    if (o.n_bits < 1 || o.n_bits > 64) return fail(DBG_ERR_ARG, "n_bits out of range");
    if (o.contraction < DBG_CONTRACT_AUTO || o.contraction > DBG_CONTRACT_SPLIT16)
        return fail(DBG_ERR_ARG, "unknown contraction");
    if (o.contraction == DBG_CONTRACT_SPLIT16 && (o.precision != DBG_PREC_MIXED || o.kind != DBG_KIND_DENSE))
        return fail(DBG_ERR_ARG, "the split-FP16 contraction is built for the mixed dense path");
    // NaN bounds would give the reference a NaN estimate; they are refused here (the
    // FP32 tail mass clamps its argument with a NaN-dropping min)
    for (int i = 0; i < n; ++i)
        if (std::isnan(a[i]) || std::isnan(b[i])) return fail(DBG_ERR_ARG, "the bounds a and b must not be NaN");
    dbg_plan* P = new dbg_plan();
    P->o = o;
    P->n = n;
    P->nd = n - 1;
    P->mixed = (o.precision == DBG_PREC_MIXED);
    P->tsz = P->mixed ? sizeof(float) : sizeof(double);
    const int64_t skip_n = 2 * o.m - 1;
    if (o.sobol_skip >= 0) {
        P->sobol_skip = (uint64_t)o.sobol_skip;
    } else {
        int k = 0;
        while ((1ll << (k + 1)) <= skip_n + 1) ++k;
        P->sobol_skip = (uint64_t)((1ll << k) - 1);
    }
    P->state = (o.kind == DBG_KIND_SPARSEQMC) ? 1 : 0;
    int rc = DBG_OK;
    hipError_t e = hipStreamCreate(&P->stream);
    if (e == hipSuccess) e = hipEventCreate(&P->ev0);
    if (e == hipSuccess) e = hipEventCreate(&P->ev1);
    if (e != hipSuccess) rc = fail(DBG_ERR_DEVICE, std::string("stream: ") + hipGetErrorString(e));
    if (!rc) rc = plan_build(P, Sigma, lda, a, b);
    if (rc) {
        std::string msg = g_err;
        P->release();
        delete P;
        g_err = msg;
        return rc;
    }
    *plan = P;
    return DBG_OK;
}

int dbg_plan_destroy(dbg_plan* plan)
{
    if (!plan) return DBG_OK;
    (void)hipStreamSynchronize(plan->stream);
    plan->release();
    delete plan;
    return DBG_OK;
}

int dbg_plan_refactor(dbg_plan* P)
{
    g_err.clear();
    if (!P) return fail(DBG_ERR_ARG, "null plan");
    return plan_factor(P);
}

int dbg_plan_set_shifts(dbg_plan* P, const void* shifts)
{
    g_err.clear();
    if (!P || (!shifts && P->nd > 0)) return fail(DBG_ERR_ARG, "null argument");
    const size_t cnt = (size_t)P->nd * P->o.n_reps;
    if (cnt > 0) {
        const bool sob = P->o.generator == DBG_GEN_SOBOL || P->o.kind == DBG_KIND_SPARSEQMC;
        if (sob)
            HIPCHK(hipMemcpyAsync(P->shifts_u, shifts, cnt * sizeof(uint64_t), hipMemcpyHostToDevice, P->stream));
        else {
            const std::vector<uint32_t> s32 = fixed_shift32((const double*)shifts, cnt);
            HIPCHK(hipMemcpyAsync(P->shifts_f, shifts, cnt * sizeof(double), hipMemcpyHostToDevice, P->stream));
            HIPCHK(hipMemcpyAsync(P->shift32, s32.data(), cnt * sizeof(uint32_t), hipMemcpyHostToDevice, P->stream));
        }
        HIPCHK(hipStreamSynchronize(P->stream));
    }
    P->shifts_set = true;
    return DBG_OK;
}

int dbg_plan_set_sobol_directions(dbg_plan* P, const uint32_t* V)
{
    g_err.clear();
    if (!P || (!V && P->nd > 0)) return fail(DBG_ERR_ARG, "null argument");
    if (P->nd > 0) {
        HIPCHK(hipMemcpyAsync(P->V, V, (size_t)P->nd * 32 * sizeof(uint32_t), hipMemcpyHostToDevice, P->stream));
        HIPCHK(hipStreamSynchronize(P->stream));
    }
    P->dirs_set = true;
    return DBG_OK;
}

int dbg_qmc_loop(dbg_plan* P, int64_t n_pts0)
{
    g_err.clear();
    if (!P) return fail(DBG_ERR_ARG, "null plan");
    const int64_t block = (P->o.kind == DBG_KIND_SPARSEQMC) ? 0 : P->state;
    return run_loop(P, block, n_pts0);
}

int dbg_plan_sync(dbg_plan* P)
{
    if (!P) return fail(DBG_ERR_ARG, "null plan");
    HIPCHK(hipStreamSynchronize(P->stream));
    return DBG_OK;
}

int dbg_plan_last_kernel_ms(dbg_plan* P, float* ms)
{
    if (!P || !ms) return fail(DBG_ERR_ARG, "null argument");
    HIPCHK(hipEventSynchronize(P->ev1));
    HIPCHK(hipEventElapsedTime(ms, P->ev0, P->ev1));
    return DBG_OK;
}

int dbg_qmc_grad(dbg_plan* P, double* log_value, double* ga, double* gb, double* gSigma, double* ga_std,
                 double* gb_std, double* gU)
{
    g_err.clear();
    if (!P) return fail(DBG_ERR_ARG, "null plan");
    const dbg_opts& o = P->o;
    const int n = P->n;
    if (o.kind != DBG_KIND_DENSE) return fail(DBG_ERR_ARG, "the gradient is defined for the dense (QMCData) plan");
    if (P->L.on) return fail(DBG_ERR_ARG, "the gradient is defined for plans of full rank");
    if (n < 2) return fail(DBG_ERR_ARG, "the gradient needs n >= 2");
    if (P->nd > 0 && !P->shifts_set) return fail(DBG_ERR_STATE, "shifts not set");
    if (o.generator == DBG_GEN_SOBOL && !P->dirs_set) return fail(DBG_ERR_STATE, "Sobol direction integers not set");
    int rc = grad_alloc(P);
    if (rc) return rc;
    HIPCHK(hipEventRecord(P->ev0, P->stream));
    rc = run_grad(P);
    if (rc) return rc;
    HIPCHK(hipEventRecord(P->ev1, P->stream));
    auto& G = P->G;
    double norm[4];
    HIPCHK(hipMemcpyAsync(norm, G.norm, 3 * sizeof(double), hipMemcpyDeviceToHost, P->stream));
    const size_t nn = (size_t)n * n * sizeof(double);
    if (ga) HIPCHK(hipMemcpyAsync(ga, G.gab, n * sizeof(double), hipMemcpyDeviceToHost, P->stream));
    if (gb) HIPCHK(hipMemcpyAsync(gb, G.gab + n, n * sizeof(double), hipMemcpyDeviceToHost, P->stream));
    if (gSigma) HIPCHK(hipMemcpyAsync(gSigma, G.g, nn, hipMemcpyDeviceToHost, P->stream));
    if (ga_std) HIPCHK(hipMemcpyAsync(ga_std, G.acc, n * sizeof(double), hipMemcpyDeviceToHost, P->stream));
    if (gb_std) HIPCHK(hipMemcpyAsync(gb_std, G.acc + n, n * sizeof(double), hipMemcpyDeviceToHost, P->stream));
    if (gU) HIPCHK(hipMemcpyAsync(gU, G.ub, nn, hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    if (log_value) *log_value = norm[2];
    return DBG_OK;
}

int dbg_plan_lowrank_info(dbg_plan* P, int32_t* rank, int32_t* k_js, int32_t* k_0)
{
    g_err.clear();
    if (!P) return fail(DBG_ERR_ARG, "null plan");
    const auto& L = P->L;
    if (rank) *rank = L.on ? L.r : P->rank;
    if (k_0) *k_0 = L.on ? L.k0 : 0;
    if (k_js) {
        if (L.on)
            std::copy(L.kjs.begin(), L.kjs.end(), k_js);
        else
            for (int i = 0; i < P->rank; ++i) k_js[i] = i;
    }
    return DBG_OK;
}

int dbg_plan_set_grad_budget(dbg_plan* P, int64_t record_bytes)
{
    g_err.clear();
    if (!P) return fail(DBG_ERR_ARG, "null plan");
    if (record_bytes < 0) return fail(DBG_ERR_ARG, "the gradient budget is negative");
    const int64_t b = record_bytes == 0 ? kGradRecordBudget : record_bytes;
    if (P->G.budget != b) {
        HIPCHK(hipStreamSynchronize(P->stream));
        P->G.free_all();
        P->G.budget = b;
    }
    return DBG_OK;
}

int dbg_plan_grad_info(dbg_plan* P, int64_t* chunk_samples, int32_t* n_chunks, int64_t* device_bytes)
{
    g_err.clear();
    if (!P) return fail(DBG_ERR_ARG, "null plan");
    const auto& G = P->G;
    if (chunk_samples) *chunk_samples = G.ready ? G.chunk_tiles * G.tile : 0;
    if (n_chunks) *n_chunks = G.ready ? G.n_schunks : 0;
    if (device_bytes) *device_bytes = G.ready ? G.bytes : 0;
    return DBG_OK;
}

int dbg_plan_refine_info(dbg_plan* P, int32_t* tiles_per_shift, double* unrefined_share, float* ms)
{
    g_err.clear();
    if (!P) return fail(DBG_ERR_ARG, "null plan");
    auto& R = P->R;
    if (tiles_per_shift) *tiles_per_shift = P->L.on ? 0 : R.K;
    if (unrefined_share) *unrefined_share = 0.0;
    if (ms) *ms = 0.0f;
    if (R.K == 0 || P->L.on) return DBG_OK;
    std::vector<double> f(P->o.n_reps, 0.0);
    HIPCHK(hipMemcpyAsync(f.data(), R.frac, f.size() * sizeof(double), hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    if (unrefined_share) *unrefined_share = *std::max_element(f.begin(), f.end());
    if (ms) HIPCHK(hipEventElapsedTime(ms, R.ev0, R.ev1));
    return DBG_OK;
}

int dbg_plan_summary(dbg_plan* P, dbg_result* out)
{
    if (!P || !out) return fail(DBG_ERR_ARG, "null argument");
    return read_summary(P, out);
}

int dbg_qmc_pnorm(dbg_plan* P, dbg_result* out)
{
    g_err.clear();
    if (!P || !out) return fail(DBG_ERR_ARG, "null argument");
    const dbg_opts& o = P->o;
    HIPCHK(hipMemsetAsync(P->rep_m, 0, (size_t)o.n_reps * sizeof(double), P->stream));
    HIPCHK(hipMemsetAsync(P->rep_e, 0, (size_t)o.n_reps * sizeof(int), P->stream));
    int64_t n_pts = 0;
    int rounds = 0;
    dbg_result r{};
    for (;;) {
        int64_t block;
        if (o.kind == DBG_KIND_SPARSEQMC)
            block = rounds == 0 ? 0 : P->state++;
        else
            block = P->state;
        int rc = run_loop(P, block, n_pts);
        if (rc) return rc;
        rounds++;
        rc = read_summary(P, &r);
        if (rc) return rc;
        n_pts += o.m;
        const double err_acc = r.err3;
        if (rounds == 1 && ((n_pts >= o.max_pts) || (err_acc < o.max_abs_err))) break;
        if (!((n_pts < o.max_pts) && (err_acc > o.max_abs_err))) break;
        if (o.kind != DBG_KIND_SPARSEQMC) P->state++;  // next_points!
    }
    r.n_pts = n_pts;
    r.rounds = rounds;
    *out = r;
    if (isnan(r.value) && isnan(r.log_value)) return fail(DBG_ERR_NONFINITE, "result is NaN");
    return DBG_OK;
}

int dbg_plan_get_factor(dbg_plan* P, double* U, int32_t* perm, double* a, double* b, double* scale,
                        int32_t* rank)
{
    g_err.clear();
    if (!P) return fail(DBG_ERR_ARG, "null plan");
    const int n = P->n;
    HIPCHK(hipStreamSynchronize(P->stream));
    if (U) HIPCHK(hipMemcpy(U, P->F.M, (size_t)n * n * sizeof(double), hipMemcpyDeviceToHost));
    if (perm) HIPCHK(hipMemcpy(perm, P->F.perm, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    // rank-deficient plans: the bounds of the reordered columns
    if (a) HIPCHK(hipMemcpy(a, P->L.on ? P->L.a : P->F.a, n * sizeof(double), hipMemcpyDeviceToHost));
    if (b) HIPCHK(hipMemcpy(b, P->L.on ? P->L.b : P->F.b, n * sizeof(double), hipMemcpyDeviceToHost));
    if (scale) HIPCHK(hipMemcpy(scale, P->F.scale, n * sizeof(double), hipMemcpyDeviceToHost));
    if (rank) *rank = P->rank;
    return DBG_OK;
}

int dbg_plan_get_reps(dbg_plan* P, double* rep_mant, int32_t* rep_exp)
{
    g_err.clear();
    if (!P) return fail(DBG_ERR_ARG, "null plan");
    HIPCHK(hipStreamSynchronize(P->stream));
    if (rep_mant) HIPCHK(hipMemcpy(rep_mant, P->rep_m, P->o.n_reps * sizeof(double), hipMemcpyDeviceToHost));
    if (rep_exp) HIPCHK(hipMemcpy(rep_exp, P->rep_e, P->o.n_reps * sizeof(int32_t), hipMemcpyDeviceToHost));
    return DBG_OK;
}

int dbg_plan_get_state(dbg_plan* P, int64_t* state)
{
    if (!P || !state) return fail(DBG_ERR_ARG, "null argument");
    *state = P->state;
    return DBG_OK;
}

int dbg_plan_set_state(dbg_plan* P, int64_t state)
{
    if (!P || state < 0) return fail(DBG_ERR_ARG, "invalid state");
    P->state = state;
    return DBG_OK;
}

}  // extern "C"

// ===========================================================================
// Batched problems of uneven dimension
// ===========================================================================
struct dbg_batch {
    dbg_opts o{};
    int K = 0;
    bool mixed = false;
    size_t tsz = 8;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    BatchArgs B{};
    std::vector<void*> owned;
    void* Uk = nullptr;
    void* Ucol = nullptr;
    void* uc16 = nullptr;      // mixed: the split-FP16 factors of the matrix-core contraction
    void* Y = nullptr;
    double* part_m = nullptr;
    int* part_e = nullptr;
    int2* sample_items = nullptr;
    int64_t n_sample = 0;
    int grid = 0;              // workspace slots: the largest resident grid over the widths
    int grid_w[3] = {0, 0, 0};  // resident grid of the integration kernel at width 8, 16, 32
    int W = 16;
    float tune_ms[3] = {0, 0, 0};

    void release()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* p : owned) (void)hipFree(p);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
    template <typename P>
    int alloc(P** p, size_t bytes)
    {
        HIPCHK(dalloc(p, bytes));
        owned.push_back((void*)*p);
        return DBG_OK;
    }
};

namespace {

// The queue loop's status word after a launch has finished (nonzero: a
// workgroup was handed an item it had already taken, so it stopped).
int batch_status(dbg_batch* Bt)
{
    int st = 0;
    HIPCHK(hipMemcpyAsync(&st, Bt->B.status, sizeof(int), hipMemcpyDeviceToHost, Bt->stream));
    HIPCHK(hipStreamSynchronize(Bt->stream));
    if (st != 0) return fail(DBG_ERR_DEVICE, "batched work queue: a workgroup fetched an item twice");
    return DBG_OK;
}

// Extents of the batch's device buffers (elements), for the host-side check
// of every offset the kernels derive from the problem table and the items.
struct BatchExtents {
    int64_t sig, vec, uk, sh, part, y_slot, nq, uc;
    int K, R, dmax, tile;
    int64_t tpr;
};

// Every access the batched kernels make, bounded on the host before upload:
// the factor work matrix and its column-major copy (d x d at sig_off), the
// packed rows (d - 1 rows of ldu; the contraction reads W <= 32 entries from
// column g0 + 1 with g0 <= d - 2, so ldu >= d + 31), the per-index vectors
// (d at vec_off), the shifts ((d - 1) x R at shift_off), the tile partials
// (tpr x R at part_off), the per-workgroup y slot ((d - 1) x tile, and for
// the split y layout whole k-groups of the blocks of 32 dimensions), the
// split-FP16 factor (d + 32 columns of ucs halves at uc_off, ucs covering every
// row < d in chunks of 32) and the lattice table (d - 1 entries); each item
// names a problem and a tile of it.
int batch_validate(const std::vector<BatchProblem>& probs, const std::vector<int2>& items,
                   const std::vector<int2>& sample, const BatchExtents& X)
{
    auto bad = [](const std::string& what) { return fail(DBG_ERR_ARG, "batched layout: " + what); };
    for (int k = 0; k < X.K; ++k) {
        const BatchProblem& P = probs[k];
        const int64_t d = P.d, nd = std::max<int64_t>(d - 1, 0);
        if (d < 1 || d > kBatchMaxDim || d > X.dmax) return bad("dimension out of range");
        if (P.sig_off < 0 || P.sig_off + d * d > X.sig || P.ucol_off != P.sig_off) return bad("factor extent");
        if (P.vec_off < 0 || P.vec_off + d > X.vec) return bad("vector extent");
        if (P.ldu < d + 31) return bad("packed row stride");
        if (P.uk_off < 0 || P.uk_off + nd * P.ldu > X.uk) return bad("packed rows extent");
        if (P.shift_off < 0 || P.shift_off + nd * X.R > X.sh) return bad("shift extent");
        if (P.tiles_per_rep != X.tpr || P.part_off < 0 || P.part_off + X.tpr * X.R > X.part)
            return bad("partials extent");
        if (nd * X.tile > X.y_slot || nd > X.nq) return bad("workspace extent");
        if ((nd + 31) / 32 * 32 * X.tile > X.y_slot) return bad("split workspace extent");
        if (X.uc > 0 && (P.ucs != (d + 31) / 32 * 64 || P.uc_off < 0 || P.uc_off + (d + 32) * P.ucs > X.uc))
            return bad("split factor extent");
    }
    for (const std::vector<int2>* v : {&items, &sample})
        for (const int2& it : *v)
            if (it.x < 0 || it.x >= X.K || it.y < 0 || (int64_t)it.y >= X.tpr * X.R) return bad("work item");
    if ((int64_t)items.size() != (int64_t)X.K * X.tpr * X.R) return bad("item count");
    return DBG_OK;
}

// The whole job as one graph: factor, (mixed) split factors, queue reset,
// integration at width Bt->W, reductions.  Replaces an earlier graph.
int batch_capture(dbg_batch* Bt)
{
    const BatchArgs& B = Bt->B;
    hipStream_t s = Bt->stream;
    const int wsel = Bt->W == 8 ? 0 : (Bt->W == 16 ? 1 : 2);
    if (Bt->exec) {
        HIPCHK(hipGraphExecDestroy(Bt->exec));
        Bt->exec = nullptr;
    }
    if (Bt->graph) {
        HIPCHK(hipGraphDestroy(Bt->graph));
        Bt->graph = nullptr;
    }
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t e = batch_factor(B, Bt->mixed, Bt->Uk, Bt->Ucol, s);
    if (e == hipSuccess && Bt->mixed) e = batch_pack_split(B, Bt->Ucol, Bt->uc16, s);
    if (e == hipSuccess) e = queue_reset(B, s);
    if (e == hipSuccess)
        e = batch_integrate(B, Bt->mixed, Bt->W, Bt->grid_w[wsel], Bt->Uk, Bt->Ucol, Bt->uc16, Bt->Y, Bt->part_m,
                            Bt->part_e, s);
    if (e == hipSuccess) e = batch_reduce(B, Bt->part_m, Bt->part_e, s);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(s, &g);
    Bt->graph = g;
    HIPCHK(e);
    HIPCHK(e2);
    HIPCHK(hipGraphInstantiate(&Bt->exec, Bt->graph, nullptr, nullptr, 0));
    HIPCHK(hipStreamSynchronize(s));
    return DBG_OK;
}

int batch_build(dbg_batch* Bt, const int32_t* dims, const double* sigmas, const double* a,
                const double* b, const double* shifts)
{
    const dbg_opts& o = Bt->o;
    const int K = Bt->K, R = o.n_reps;
    const size_t T = Bt->tsz;
    hipStream_t s = Bt->stream;
    BatchArgs& B = Bt->B;
    std::vector<BatchProblem> probs(K);
    int64_t sig = 0, vec = 0, uk = 0, sh = 0, part = 0, ucn = 0;
    int dmax = 1;
    const int tile = batch_tile();
    const int64_t tpr = (o.m + tile - 1) / tile;
    for (int k = 0; k < K; ++k) {
        const int d = dims[k];
        if (d < 1 || d > kBatchMaxDim)
            return fail(DBG_ERR_ARG, "batched dimensions must be in [1, " + std::to_string(kBatchMaxDim) + "]");
        if (d > o.chol_block_size2 || d > o.chol_block_size3 || d > o.chol_block_size4)
            return fail(DBG_ERR_ARG, "batched problems must fit one pivot window (d <= chol_block_size2)");
        BatchProblem& P = probs[k];
        P.d = d;
        P.tiles_per_rep = (int32_t)tpr;
        P.ldu = d + 32;
        P.lower = 1;
        for (int i = 0; i < d; ++i) P.lower = P.lower && std::isinf(a[vec + i]) && a[vec + i] < 0;
        P.ucs = (d + 31) / 32 * 64;
        P.uc_off = ucn;
        P.pad = 0;
        ucn += (int64_t)(d + 32) * P.ucs;
        P.sig_off = sig;
        P.ucol_off = sig;
        P.vec_off = vec;
        P.uk_off = uk;
        P.shift_off = sh;
        P.part_off = part;
        sig += (int64_t)d * d;
        vec += d;
        uk += (int64_t)std::max(d - 1, 0) * P.ldu;
        sh += (int64_t)(d - 1) * R;
        part += tpr * R;
        dmax = std::max(dmax, d);
    }
    // work items, largest dimension first (stable in the input order)
    std::vector<int> order(K);
    for (int k = 0; k < K; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return dims[x] > dims[y]; });
    std::vector<int2> items;
    items.reserve((size_t)K * tpr * R);
    for (int k : order)
        for (int64_t lt = 0; lt < tpr * R; ++lt) items.push_back(make_int2(k, (int)lt));
    std::vector<int2> sample;
    for (size_t i = 0; i < items.size(); i += 16) sample.push_back(items[i]);
    {
        BatchExtents X{};
        X.sig = sig;
        X.vec = vec;
        X.uk = uk;
        X.sh = sh;
        X.part = part;
        // whole blocks of 32 dimensions: the split y layout of the matrix-core contraction
        X.y_slot = (int64_t)((std::max(dmax - 1, 1) + 31) / 32 * 32) * tile;
        X.uc = Bt->mixed ? ucn : 0;
        X.nq = std::max(dmax - 1, 1);
        X.K = K;
        X.R = R;
        X.dmax = dmax;
        X.tile = tile;
        X.tpr = tpr;
        const int vrc = batch_validate(probs, items, sample, X);
        if (vrc) return vrc;
    }

    BatchProblem* dprobs;
    int rc;
#define BALLOC(p, bytes)                       \
    do {                                       \
        rc = Bt->alloc(&(p), (bytes));         \
        if (rc) return rc;                     \
    } while (0)
    BALLOC(dprobs, sizeof(BatchProblem) * K);
    double *sig_in, *a_in, *b_in, *sh_d, *q_d;
    BALLOC(sig_in, sig * sizeof(double));
    BALLOC(a_in, vec * sizeof(double));
    BALLOC(b_in, vec * sizeof(double));
    BALLOC(sh_d, std::max<int64_t>(sh, 1) * sizeof(double));
    BALLOC(B.M, sig * sizeof(double));
    BALLOC(B.a, vec * sizeof(double));
    BALLOC(B.b, vec * sizeof(double));
    BALLOC(B.scale, vec * sizeof(double));
    BALLOC(B.svec, vec * sizeof(double));
    BALLOC(B.perm, vec * sizeof(int32_t));
    BALLOC(B.rank, K * sizeof(int32_t));
    BALLOC(B.ak, vec * T);
    BALLOC(B.bk, vec * T);
    BALLOC(B.invk, vec * T);
    BALLOC(B.first, (size_t)8 * K * sizeof(double));
    BALLOC(Bt->Uk, std::max<int64_t>(uk, 1) * T);
    BALLOC(Bt->Ucol, sig * T);
    if (Bt->mixed) BALLOC(Bt->uc16, std::max<int64_t>(ucn, 1) * sizeof(_Float16));
    int2* items_d;
    BALLOC(items_d, items.size() * sizeof(int2));
    BALLOC(Bt->sample_items, std::max<size_t>(sample.size(), 1) * sizeof(int2));
    BALLOC(B.counter, sizeof(unsigned long long));
    BALLOC(B.status, sizeof(int));
    const std::vector<double> qh = sqrt_prime_fracs(dmax - 1);
    const std::vector<uint64_t> q64h = fixed_q64(qh);
    const std::vector<uint32_t> s32h = fixed_shift32(shifts, (size_t)sh);
    BALLOC(q_d, qh.size() * sizeof(double));
    uint64_t* q64_d;
    uint32_t* s32_d;
    BALLOC(q64_d, q64h.size() * sizeof(uint64_t));
    BALLOC(s32_d, std::max<int64_t>(sh, 1) * sizeof(uint32_t));
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    // the persistent integration launch fills the device: as many workgroups as
    // are resident at once for each width (registers, no LDS beyond the queue word)
    const int widths[3] = {8, 16, 32};
    Bt->grid = 0;
    for (int w = 0; w < 3; ++w) {
        const int occ = std::max(1, std::min(8, batch_occupancy(Bt->mixed, widths[w])));
        Bt->grid_w[w] = ncu * occ;
        Bt->grid = std::max(Bt->grid, Bt->grid_w[w]);
    }
    B.y_slot = (int64_t)((std::max(dmax - 1, 1) + 31) / 32 * 32) * tile;
    BALLOC(Bt->Y, (size_t)Bt->grid * B.y_slot * T);
    BALLOC(Bt->part_m, std::max<int64_t>(part, 1) * sizeof(double));
    BALLOC(Bt->part_e, std::max<int64_t>(part, 1) * sizeof(int));
    BALLOC(B.rep_m, (size_t)K * R * sizeof(double));
    BALLOC(B.rep_e, (size_t)K * R * sizeof(int));
    BALLOC(B.out, (size_t)8 * K * sizeof(double));
#undef BALLOC
    HIPCHK(hipMemcpyAsync(dprobs, probs.data(), sizeof(BatchProblem) * K, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(sig_in, sigmas, sig * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(a_in, a, vec * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b_in, b, vec * sizeof(double), hipMemcpyHostToDevice, s));
    if (sh > 0) HIPCHK(hipMemcpyAsync(sh_d, shifts, sh * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(items_d, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice, s));
    if (!sample.empty())
        HIPCHK(hipMemcpyAsync(Bt->sample_items, sample.data(), sample.size() * sizeof(int2), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(q_d, qh.data(), qh.size() * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(q64_d, q64h.data(), q64h.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    if (sh > 0) HIPCHK(hipMemcpyAsync(s32_d, s32h.data(), sh * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    B.K = K;
    B.n_reps = R;
    B.m = o.m;
    B.bs1 = o.chol_block_size;
    B.probs = dprobs;
    B.sig_in = sig_in;
    B.a_in = a_in;
    B.b_in = b_in;
    B.items = items_d;
    B.n_items = (int64_t)items.size();
    B.q = q_d;
    B.shifts = sh_d;
    B.q64 = q64_d;
    B.shift32 = s32_d;
    Bt->n_sample = (int64_t)sample.size();

    // ---- autotune the dimension-block width on a sample of the work items
    HIPCHK(batch_factor(B, Bt->mixed, Bt->Uk, Bt->Ucol, s));
    if (Bt->mixed) HIPCHK(batch_pack_split(B, Bt->Ucol, Bt->uc16, s));
    float best = 1e30f;
    BatchArgs Bs = B;
    Bs.items = Bt->sample_items;
    Bs.n_items = Bt->n_sample;
    for (int w = 0; w < 3; ++w) {
        float msum = 0.f;
        for (int rep = 0; rep < 2; ++rep) {  // first pass warms up
            HIPCHK(queue_reset(B, s));
            HIPCHK(hipEventRecord(Bt->ev0, s));
            HIPCHK(batch_integrate(Bs, Bt->mixed, widths[w], Bt->grid_w[w], Bt->Uk, Bt->Ucol, Bt->uc16, Bt->Y,
                                   Bt->part_m, Bt->part_e, s));
            HIPCHK(hipEventRecord(Bt->ev1, s));
            HIPCHK(hipEventSynchronize(Bt->ev1));
            rc = batch_status(Bt);
            if (rc) return rc;
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, Bt->ev0, Bt->ev1));
            if (rep == 1) msum = ms;
        }
        Bt->tune_ms[w] = msum;
        if (msum < best) {
            best = msum;
            Bt->W = widths[w];
        }
    }
    return batch_capture(Bt);
}

}  // namespace

extern "C" {

int dbg_batch_create(dbg_batch** out, const dbg_opts* opts, int32_t K, const int32_t* dims,
                     const double* sigmas, const double* a, const double* b, const double* shifts)
{
    g_err.clear();
    if (!out || !opts || !dims || !sigmas || !a || !b || K < 1) return fail(DBG_ERR_ARG, "invalid argument");
    *out = nullptr;
    for (int32_t k = 0; k < K; ++k)
        if (dims[k] > 1 && !shifts) return fail(DBG_ERR_ARG, "shifts must not be NULL when a dimension exceeds 1");
    if (opts->m < 1 || opts->n_reps < 1) return fail(DBG_ERR_ARG, "m and n_reps must be positive");
    if (opts->generator != DBG_GEN_RICHTMYER || opts->kind != DBG_KIND_DENSE || opts->phi_form != DBG_PHI_ERF)
        return fail(DBG_ERR_ARG, "batched problems use the dense Richtmyer erf-form path");
    dbg_batch* Bt = new dbg_batch();
    Bt->o = *opts;
    Bt->K = K;
    Bt->mixed = opts->precision == DBG_PREC_MIXED;
    Bt->tsz = Bt->mixed ? sizeof(float) : sizeof(double);
    int rc = DBG_OK;
    hipError_t e = hipStreamCreate(&Bt->stream);
    if (e == hipSuccess) e = hipEventCreate(&Bt->ev0);
    if (e == hipSuccess) e = hipEventCreate(&Bt->ev1);
    if (e != hipSuccess) rc = fail(DBG_ERR_DEVICE, std::string("stream: ") + hipGetErrorString(e));
    if (!rc) rc = batch_build(Bt, dims, sigmas, a, b, shifts);
    if (rc) {
        std::string msg = g_err;
        Bt->release();
        delete Bt;
        g_err = msg;
        return rc;
    }
    *out = Bt;
    return DBG_OK;
}

int dbg_batch_run(dbg_batch* Bt, double* results)
{
    g_err.clear();
    if (!Bt) return fail(DBG_ERR_ARG, "null batch");
    HIPCHK(hipEventRecord(Bt->ev0, Bt->stream));
    HIPCHK(hipGraphLaunch(Bt->exec, Bt->stream));
    HIPCHK(hipEventRecord(Bt->ev1, Bt->stream));
    if (results)
        HIPCHK(hipMemcpyAsync(results, Bt->B.out, (size_t)8 * Bt->K * sizeof(double), hipMemcpyDeviceToHost,
                              Bt->stream));
    HIPCHK(hipStreamSynchronize(Bt->stream));
    return batch_status(Bt);
}

int dbg_batch_last_ms(dbg_batch* Bt, float* ms)
{
    if (!Bt || !ms) return fail(DBG_ERR_ARG, "null argument");
    HIPCHK(hipEventElapsedTime(ms, Bt->ev0, Bt->ev1));
    return DBG_OK;
}

int dbg_batch_set_width(dbg_batch* Bt, int32_t width)
{
    g_err.clear();
    if (!Bt) return fail(DBG_ERR_ARG, "null batch");
    if (width != 8 && width != 16 && width != 32) return fail(DBG_ERR_ARG, "batch width must be 8, 16 or 32");
    HIPCHK(hipStreamSynchronize(Bt->stream));
    Bt->W = width;
    return batch_capture(Bt);
}

int dbg_batch_tuning(dbg_batch* Bt, int32_t* width, float* ms_w8_w16_w32)
{
    if (!Bt) return fail(DBG_ERR_ARG, "null batch");
    if (width) *width = Bt->W;
    if (ms_w8_w16_w32)
        for (int i = 0; i < 3; ++i) ms_w8_w16_w32[i] = Bt->tune_ms[i];
    return DBG_OK;
}

int dbg_batch_destroy(dbg_batch* Bt)
{
    if (!Bt) return DBG_OK;
    (void)hipStreamSynchronize(Bt->stream);
    Bt->release();
    delete Bt;
    return DBG_OK;
}

}  // extern "C"
