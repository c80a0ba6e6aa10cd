// Internal (non-ABI) declarations shared by the kernel translation units and
// the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dbg {

enum { kGenRichtmyer = 0, kGenSobol = 1 };
// kPhiTailLower: the tail form for a lower orthant (every lower bound -inf),
// an internal specialisation of the mixed walk that forms only the upper
// bound's tail mass (the lower one is exactly 0); same values as kPhiTail
enum { kPhiErf = 0, kPhiErfc = 1, kPhiTail = 2, kPhiTailLower = 3 };

// Where the points of one round (one block of m points per shift) come from.
struct PointArgs {
    int nd;                    // sampled dimensions (n - 1)
    const double* q;           // Richtmyer: frac(sqrt(p_g)), nd entries
    const double* shifts_f;    // Richtmyer: nd x n_reps shifts in [0, 1)
    int64_t n_base;            // Richtmyer: point j uses n_base + j (j 1-based)
    const uint32_t* V;         // Sobol: nd x 32 direction integers
    const uint64_t* shifts_u;  // Sobol: nd x n_reps digital shifts
    uint64_t sobol_first;      // Sobol: point j uses sequence index sobol_first + j
    uint64_t mask;             // Sobol: 2^L - 1
    double invscale;           // Sobol: 2^-L
    int sobol_bits;            // Sobol: L
    const uint64_t* q64;       // mixed Richtmyer: floor(q_g 2^64)
    const uint32_t* shift32;   // mixed Richtmyer: floor(shift 2^32), nd x n_reps
};

// Arguments of the dense integration kernels.
struct DenseArgs {
    int n;                     // dimension d
    int64_t m;                 // points per shift in this round
    int n_reps;
    int64_t tiles_per_rep;     // ceil(m / tile)
    int64_t n_tiles;           // tiles_per_rep * n_reps
    const void* Uk;            // packed strict-upper rows (T), Uk[l*ldu + j] = U[l, j] for j > l else 0
    int64_t ldu;
    const void* Ucol;          // the factor column-major (T), Ucol[j*ldc + l] = U[l, j]
    int64_t ldc;
    const void* ak;            // per target j: a_j (T)
    const void* bk;            // per target j: b_j (T)
    const void* invk;          // per target j: 1 / (U_jj * sqrt 2) (T)
    double first[6];           // first dimension: c1, dc1, ea1, eb1, d1, mode1 (FP64)
    PointArgs pts;
    void* Y;                   // workspace: one slot of nd x tile samples per resident block
    int64_t y_slot;            // elements per slot
    double* part_m;            // per-tile partial sums (scaled)
    int* part_e;
    // optional list of the tiles to run (FP64 integration kernels): entry i of
    // n_tiles names tile tile_list[i] of the round, entries < 0 are skipped
    const int64_t* tile_list;
    // rank-deficient plans (qmc_lowrank.hip): rank + 1 column-block starts
    // of the reordered factor (Ucol, ldc) and the rank r (r - 1 sampled dimensions)
    const int32_t* lr_beg;
    int lr_rank;
    // adjoint support (matrix-core kernels only)
    int mode;                  // kModeIntegrate, kModeRecord or kModeReverse
    // record / reverse: the launch runs tiles tile0 .. tile0 + n_tiles - 1 of
    // the round (one chunk of samples); records, weights and partial sums
    // are indexed from the chunk's first tile
    int64_t tile0;
    int64_t rs;                // row stride of the records: chunk tiles * tile samples
    void* srec;                // record: s_{g+1} per (g, sample); reverse: replaced by sbar_{g+1}
    void* yrec;                // record: y_g per (g, sample) (the workspace itself in record mode)
    void* wm;                  // record: per-sample weight mantissa (T)
    int* we;                   // record: per-sample weight exponent
    const double* norm;        // reverse: {E, 1 / sum(p 2^-E)}
    double* ga;                // reverse: d log P / d a', b', U_rr accumulators (n each)
    double* gb;
    double* gd;
    int a_inf;                 // every lower bound is -inf (mixed walk: kPhiTailLower; reverse: d log P / d a' is 0)
};
enum { kModeIntegrate = 0, kModeRecord = 1, kModeReverse = 2 };

// Arguments of the sparse (CSC) integration kernel.
struct SparseArgs {
    int n;
    int64_t m;
    int n_reps;
    int64_t tiles_per_rep;
    int64_t n_tiles;
    const int64_t* colptr;     // n + 1, 0-based
    const int32_t* rowval;     // 0-based rows
    const void* nzval;         // T
    const void* ak;
    const void* bk;
    const void* invk;
    double first[6];
    int form;                  // kPhiErf or kPhiTail
    int tent;                  // 1: src/qmc.jl form, 0: SparseQMC (no tent)
    PointArgs pts;
    void* Y;
    int64_t y_slot;
    double* part_m;
    int* part_e;
};

hipError_t launch_dense_valu(const DenseArgs& A, bool mixed, int gen, int form, int grid,
                             hipStream_t s);
int dense_valu_tile();
int dense_valu_width();
hipError_t launch_dense_mfma(const DenseArgs& A, bool mixed, int gen, int form, int grid,
                             hipStream_t s);
// the rank-deficient route (qmc_lowrank.hip)
int lowrank_tile();
hipError_t lowrank_fill(double* U, int n, int r, const double* Sig, const int32_t* perm, const double* scale,
                        hipStream_t s);
hipError_t lowrank_bounds(const double* a, const double* b, const int32_t* perm, int n, double* ap, double* bp,
                          hipStream_t s);
hipError_t lowrank_prep(const double* U, int n, const int32_t* beg, int r, const double* a, const double* b,
                        double* inv, double* first, int phi_form, hipStream_t s);
hipError_t launch_lowrank(const DenseArgs& A, int gen, int form, int grid, hipStream_t s);
// record (forward keeping y, s and the weights) / reverse sweep, tail form
hipError_t launch_dense_adjoint(const DenseArgs& A, bool mixed, int gen, int grid, hipStream_t s);
hipError_t pack_flip(const double* U, int n, bool mixed, void* Ufk, int64_t ldu, void* Ufc, hipStream_t s);
hipError_t grad_norm(const double* pm, const int* pe, int64_t n_tiles, double total, double* norm, hipStream_t s);
// folds one chunk of samples into the running gradient: norm (E, 1/S, ., S of
// the chunk) and run (E, S of the chunks before) give fold = {alpha, beta},
// the weights of the running sums and of the chunk's; with last != 0 also
// log P of all chunks into norm[2]
hipError_t grad_fold_norm(double* norm, double* run, double* fold, int first, int last, double total,
                          hipStream_t s);
// acc[i] = fold[0] acc[i] + fold[1] cacc[i], i < len
hipError_t grad_fold_acc(double* acc, const double* cacc, int len, const double* fold, hipStream_t s);
// Ub[l, g+1] = fold[0] Ub[l, g+1] + fold[1] sum_s Y[l, s] S[g, s] over the
// first kmax samples of rows of stride rs (l <= g < nd).
// split16: the mixed product on FP16 matrix cores (high/low split, per-row scales of S)
hipError_t ubar_product(const void* Y, const void* S, int64_t rs, int64_t kmax, int nd, bool mixed,
                        const int2* pairs, int n_pairs, int n_chunks, int64_t kchunk, void* part, double* Ub, int n,
                        const double* fold, hipStream_t s, bool split16);
int ubar_tile(bool split16);        // output tile of the Ubar product
int ubar_wgs_per_cu(bool split16);  // its resident workgroups per CU
// step 0: Ubar diagonal; 1: Phi(X); 2: C_bar -> G, ga, gb
hipError_t sigma_chain(double* Ub, int n, const double* gd, double* X, hipStream_t s, int step,
                       const double* Cb, const int32_t* perm, const double* scale, const double* Sig,
                       const double* a_std, const double* b_std, const double* ga_std, const double* gb_std,
                       double* G, double* ga, double* gb);
// the rocBLAS handle of this host thread and device (rocblas_handle), stream set to s
void* blas_handle_void(hipStream_t s);
int dense_mfma_tile(bool mixed);
int dense_mfma_panel(bool mixed);
// the FP64 re-evaluation of a mixed round's heaviest tiles: the FP64
// matrix-core kernel with 128-sample tiles and 128-target panels, tail form,
// on a tile list (latency-bound: a few tiles per shift)
int dense_mfma_refine_tile();
int dense_mfma_refine_panel();
hipError_t launch_dense_mfma_refine(const DenseArgs& A, int gen, int grid, hipStream_t s);
// mixed dense kernel with the panel contraction on FP16 matrix cores
// (split operands; Richtmyer points)
hipError_t launch_dense_split(const DenseArgs& A, int gen, const void* uc, int grid, hipStream_t s);
hipError_t pack_split16(const float* Ucol, int n, void* uc, hipStream_t s);
int64_t split16_col_stride(int n);  // halfs per target column of the split factor
int split16_cols(int n);            // target columns of the split factor
int dense_split_tile();
int64_t dense_split_slot(int nd);   // y workspace floats per workgroup slot
// dimension from which the dense path uses the matrix-core kernel
constexpr int kMfmaMinDim = 96;
// FP64 tiles re-evaluated per shift by a mixed dense plan unless its options say otherwise
constexpr int kDefaultRefineTiles = 8;
// default device memory for the adjoint's y and s records (bytes); larger
// jobs run in chunks of samples
constexpr int64_t kGradRecordBudget = int64_t(4) << 30;

// ring: 0 (y rows in the global workspace) or the LDS ring size from sparse_ring
hipError_t launch_sparse(const SparseArgs& A, bool mixed, int gen, int ring, int grid, hipStream_t s);
int sparse_tile();
int sparse_ring(bool mixed, int reach);
hipError_t sparse_grid(bool mixed, int gen, int form, int ring, int ncu, int* grid);

// ---- batched problems of uneven dimension --------------------------------
struct BatchProblem {
    int32_t d;
    int32_t tiles_per_rep;
    int64_t sig_off;    // offset of Sigma_k / the factor (d x d doubles)
    int64_t vec_off;    // offset of a, b, scale, perm, ak, bk, invk (d entries)
    int64_t uk_off;     // offset of the packed rows (d - 1 rows of ldu, working precision)
    int64_t ucol_off;   // offset of the column-major factor copy (d x d, working precision)
    int64_t shift_off;  // offset of the shifts ((d - 1) x n_reps doubles)
    int64_t part_off;   // offset of the problem's tile partials (n_reps x tiles_per_rep)
    int32_t ldu;
    int32_t lower;      // 1 when every lower bound of the problem is -inf (the mixed walk's kPhiTailLower)
    int64_t uc_off;     // mixed: offset (halves) of the split-FP16 factor, d + 32 columns of ucs halves
    int32_t ucs;        // mixed: halves per split column, 64 per chunk of 32 rows
    int32_t pad;
};

struct BatchArgs {
    int K;
    int n_reps;
    int64_t m;
    int bs1;
    const BatchProblem* probs;
    const double* sig_in;
    const double* a_in;
    const double* b_in;
    double* M;
    double* a;
    double* b;
    double* scale;
    double* svec;
    int32_t* perm;
    int32_t* rank;
    void* ak;
    void* bk;
    void* invk;
    double* first;                 // 8 per problem
    const int2* items;             // (problem, rep * tiles_per_rep + tile), largest d first
    int64_t n_items;
    unsigned long long* counter;   // work-queue head
    int* status;                   // nonzero: a workgroup fetched an item twice (queue loop stopped)
    const double* q;
    const double* shifts;
    const uint64_t* q64;           // q in 64-bit fixed point (mixed)
    const uint32_t* shift32;       // shifts in 32-bit fixed point (mixed)
    int64_t y_slot;
    double* rep_m;                 // n_reps per problem
    int* rep_e;
    double* out;                   // 8 per problem: value, log, err3, rel, mean_m, err_m, E, rank
};

constexpr int kBatchMaxDim = 512;
hipError_t batch_factor(const BatchArgs& B, bool mixed, void* Uk, void* Ucol, hipStream_t s);
int batch_occupancy(bool mixed, int W);
hipError_t batch_integrate(const BatchArgs& B, bool mixed, int W, int grid, const void* Uk,
                           const void* Ucol, const void* uc16, void* Y, double* part_m, int* part_e, hipStream_t s);
hipError_t batch_pack_split(const BatchArgs& B, const void* Ucol, void* uc16, hipStream_t s);
hipError_t batch_reduce(const BatchArgs& B, const double* part_m, const int* part_e, hipStream_t s);
hipError_t queue_reset(const BatchArgs& B, hipStream_t s);
int batch_tile();

// Per-shift reduction of the tile partials and running-mean merge
// (src/qmc.jl:434-442); then mean and 3*sd over shifts (:303-312, 814-815).
hipError_t launch_reduce_reps(const double* part_m, const int* part_e, int64_t tiles_per_rep,
                              int n_reps, int64_t n_pts0, int64_t m, int sparseqmc_merge,
                              double* rep_m, int* rep_e, double* out, hipStream_t s);

// Cholesky building blocks (FP64, column-major, lda).
struct CholArgs {
    double* M;
    int64_t lda;
    int n;
    double* a;
    double* b;
    double* scale;
    int32_t* perm;
    double* s_vec;
    double* eps0;      // device scalar
    int32_t* rank;     // device scalar (n while full rank)
};
hipError_t chol_standardize(const CholArgs& C, bool mirror, hipStream_t s);
hipError_t chol_window(const CholArgs& C, int w0, int w1, int nb, int pivot, hipStream_t s);
hipError_t chol_trsm_syrk(const CholArgs& C, int r0, int r1, int c0, int c1, hipStream_t s);
hipError_t chol_finish(const CholArgs& C, hipStream_t s);
// covariance generators (cov.hip); kinds as DBG_COV_* of the C ABI
constexpr int kCovFixedDense = 0, kCovFixedSparse = 1, kCovMatern = 2, kCovLowRank = 3;
hipError_t cov_fill(int kind, int n, double* S, int64_t ld, double p0, double p1, double p2, hipStream_t s);
int cov_lowrank(int n, int k, const double* B, const double* psi, double* S, int64_t ld, hipStream_t s);
hipError_t cov_standardize(int n, double* S, int64_t ld, double* scratch, hipStream_t s);
// adapt_low_rank! on device arrays (M n x n, lda; end: rank entries, out; tmp:
// n * n doubles of scratch; k0: one entry, out)
hipError_t adapt_low_rank(double* M, int64_t lda, int n, double* a, double* b, int32_t* perm, int rank, double eps1,
                          int32_t* end, double* tmp, int32_t* k0, hipStream_t s);
int chol_max_window();

// Post-factor preparation.
hipError_t truncate_upper(double* U, int64_t lda, int n, double min_val, hipStream_t s);
hipError_t pack_dense(const double* U, int64_t lda, int n, const double* a, const double* b,
                      bool mixed, void* Uk, int64_t ldu, void* ak, void* bk, void* invk,
                      double* first /* c1, dc1, ea1, eb1, d1, mode1 */, int phi_form,
                      hipStream_t s);
hipError_t convert_f32(const double* in, size_t count, float* out, hipStream_t s);
// FP64 re-evaluation of the heaviest tiles of a mixed round (refine.hip):
// per shift, the K tiles with the largest partial sums (sel: global mixed tile
// indices, list64: the FP64 tiles covering them, ratio per mixed tile; -1
// where fewer than K tiles are nonzero) and the share of the shift's sum the
// other tiles hold (frac); then the picked partials from the FP64 tiles.
hipError_t refine_select(const double* part_m, const int* part_e, int64_t tpr, int n_reps, int K, int ratio,
                         int64_t tpr64, int64_t* sel, int64_t* list64, double* frac, hipStream_t s);
hipError_t refine_merge(const int64_t* sel, int n_reps, int K, int ratio, int64_t tpr, int64_t tpr64,
                        const double* p64_m, const int* p64_e, double* part_m, int* part_e, hipStream_t s);
hipError_t csc_count(const double* U, int64_t lda, int n, int64_t* colcount, hipStream_t s);
hipError_t csc_fill(const double* U, int64_t lda, int n, const int64_t* colptr, bool mixed,
                    int32_t* rowval, void* nzval, hipStream_t s);
hipError_t exclusive_scan_i64(const int64_t* in, int64_t* out, int n, hipStream_t s);
hipError_t csc_reach(const int64_t* colptr, const int32_t* rowval, int n, int* out, hipStream_t s);

}  // namespace dbg
