/*
 * dbgenz.h -- C ABI of the MI355X-native Genz quasi-Monte Carlo
 * multivariate-normal probability library (libdbgenz.so).
 *
 * The entry points replace the hot path of JakobPeterlin/DBGenzPaper, a pure
 * Julia package.  Each declaration names the reference routine it stands in
 * for (paths relative to the reference repository root).  A Julia caller binds
 * them with one ccall per entry point.
 *
 * Conventions
 *   - Matrices are column-major with an explicit leading dimension, like a
 *     Julia Matrix{Float64}.  Permutations are 0-based (Julia adds 1).
 *   - Host pointers unless the name ends in _dev.
 *   - Every function returns a status code (DBG_OK on success);
 *     dbg_last_error() returns a message for the calling thread.
 *   - A plan owns its device buffers and its HIP stream.  Calls on one plan
 *     are ordered; distinct plans are independent and may be used from
 *     different host threads.  No call keeps global mutable state.
 */
#ifndef DBGENZ_H
#define DBGENZ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define DBG_OK 0
#define DBG_ERR_ARG 1        /* invalid argument (size, option, null pointer) */
#define DBG_ERR_RANK 2       /* factor not of full rank: QMCData's ArgumentError (src/qmc.jl:195-197) */
#define DBG_ERR_DEVICE 3     /* HIP runtime error or out of device memory */
#define DBG_ERR_NONFINITE 4  /* result is NaN */
#define DBG_ERR_STATE 5      /* call out of order (e.g. no shifts set) */

/* precision of the integration (the Julia type parameter T) */
#define DBG_PREC_FP64 0   /* T = Float64: the reference arithmetic in FP64 */
#define DBG_PREC_MIXED 1  /* T = Float32: FP32 contraction and special functions,
                             32-bit fixed-point points, FP64 factor and reductions */
/* point set (qmc_type) */
#define DBG_GEN_RICHTMYER 0
#define DBG_GEN_SOBOL 1
/* form of the normal CDF in the loop */
#define DBG_PHI_ERF 0   /* QMCData (src/qmc.jl:405-429) */
#define DBG_PHI_ERFC 1  /* QMCData_Distributions (src/qmc.jl:579-593) */
#define DBG_PHI_TAIL 2  /* no reference counterpart: complementary tail masses without the
                           [-9, 9] clamp, Phi^{-1} from the nearer tail (erfcinv); keeps
                           problems whose per-dimension masses fall below 1e-37 finite.
                           The FP32 mixed path always uses this form. */
/* which data structure / loop */
#define DBG_KIND_DENSE 0      /* QMCData: pivoted factor, dense loop */
#define DBG_KIND_SPARSE 1     /* QMCDataSparse: unpivoted factor, CSC loop */
#define DBG_KIND_SPARSEQMC 2  /* SparseQMC.QMCDataSparse: Sobol, no tent map */
/* arithmetic of the dense panel contraction in the mixed mode (no reference
 * counterpart: the reference's Float32 mode calls BLAS sgemm) */
#define DBG_CONTRACT_AUTO 0     /* the library's choice: SPLIT16 where it is built, else FP32 */
#define DBG_CONTRACT_FP32 1     /* FP32 matrix cores (v_mfma_f32_16x16x4f32) */
#define DBG_CONTRACT_SPLIT16 2  /* FP16 matrix cores, operands split into high and low
                                   FP16 parts, three products per pair, FP32 accumulation
                                   (FP32-level accuracy) */

/* Options: QMCOpts{T} (src/qmc.jl:15-27, defaults of QMC_opts :111-125) plus
 * the choices the Julia constructors take as arguments. */
typedef struct dbg_opts {
    int32_t chol_block_size;  /* bs1 of cholesky_genz!/cholesky_classic! */
    int32_t chol_block_size2; /* bs2: the pivot window */
    int32_t chol_block_size3; /* bs3 (Julia default 2^10) */
    int32_t chol_block_size4; /* bs4 (Julia default 2^11) */
    int64_t m;                /* points per shift and round */
    int32_t block_size_i;     /* accepted for API parity; the device picks its own tiling */
    int32_t block_size_i2;    /* accepted for API parity */
    int32_t block_size_j;     /* accepted for API parity */
    int32_t n_blocks;         /* accepted for API parity (CPU thread slots) */
    int32_t n_reps;           /* number of random shifts */
    int32_t n_bits;           /* Sobol bits L */
    int64_t max_pts;          /* adaptive stop: total points per shift */
    double max_abs_err;       /* adaptive stop: 3 * standard error */
    int32_t precision;        /* DBG_PREC_* */
    int32_t generator;        /* DBG_GEN_* */
    int32_t phi_form;         /* DBG_PHI_* */
    int32_t kind;             /* DBG_KIND_* */
    int64_t sobol_skip;       /* next! calls before the first Sobol point; < 0: as skip(s, 2m-1) */
    int32_t contraction;      /* DBG_CONTRACT_* (mixed dense path) */
    int32_t refine_tiles;     /* mixed dense path: after each round, the refine_tiles tiles of 256
                                 samples with the largest sums of every shift are evaluated again in
                                 FP64 (the plain FP64 path's arithmetic) and replace their FP32 sums;
                                 0: none; < 0: the library's default (8).  No reference counterpart:
                                 the reference's Float32 mode has no FP64 step in the loop */
} dbg_opts;

/* Result of qmc_pnorm!: (value, err, n_pts) of src/qmc.jl:829,849 plus the
 * log of the value, which stays finite when P is below the Float64 range. */
typedef struct dbg_result {
    double value;     /* P, 0 if it underflows Float64 */
    double log_value; /* log P */
    double err3;      /* 3 * standard error over the shifts (src/qmc.jl:815) */
    double rel_err3;  /* err3 / value, finite when both underflow */
    int64_t n_pts;    /* points per shift used */
    int32_t rounds;   /* qmc_loop! calls made */
} dbg_result;

typedef struct dbg_plan dbg_plan;

/* message of the last failed call on this thread ("" if none) */
const char* dbg_last_error(void);
/* ABI version, bumped on incompatible changes */
int32_t dbg_abi_version(void);
/* QMC_opts(T) defaults (src/qmc.jl:111-125) */
void dbg_opts_default(dbg_opts* opts, int32_t precision);

/* cholesky_genz!(M, a, b, bs1, bs2, bs3, bs4)  src/cholesky.jl:445-628
 * In place on host arrays: M (n x n, lda) receives the standardized pivoted
 * factor U in its upper triangle (strict lower triangle zeroed), a and b are
 * standardized and permuted.  perm (0-based), scale (scale_vec) and rank are
 * outputs.  Runs on the device in FP64. */
int dbg_chol_genz_f64(double* M, int64_t lda, int32_t n, double* a, double* b, int32_t bs1,
                      int32_t bs2, int32_t bs3, int32_t bs4, int32_t* perm, double* scale,
                      int32_t* rank);

/* cholesky_rowmax!(M, a, b, bs1, bs2, bs3, bs4)  src/cholesky.jl:266-432
 * The blocked loop nest of cholesky_genz! with diagonal pivoting: at each
 * step the largest remaining Schur diagonal of the pivot window (first index
 * on ties) becomes the pivot; the factorization stops (rank = pivots done)
 * where that diagonal is not positive.  In place on host arrays: M receives
 * the standardized factor U in its upper triangle (strict lower triangle
 * zeroed); a and b are standardized but, as in the reference, not permuted.
 * perm (0-based), scale (scale_vec) and rank are outputs.  FP64 on the
 * device. */
int dbg_chol_rowmax_f64(double* M, int64_t lda, int32_t n, double* a, double* b, int32_t bs1,
                        int32_t bs2, int32_t bs3, int32_t bs4, int32_t* perm, double* scale,
                        int32_t* rank);

/* adapt_low_rank!(M, a, b, rank, perm, eps1)  src/cholesky.jl:20-123
 * Reorders the dependent columns rank .. n-1 of a rank-deficient factor M
 * (n x n, lda) in place: each column whose last entry above row rank with
 * |value| > eps1 sits in row i_j is scaled so that entry is 1 (a, b scaled
 * too and exchanged when it was negative) and moved to the end of the block
 * of level i_j; columns with no such entry go to the end.  a, b and perm
 * (0-based here) move with their columns.  k_js (rank entries, out) receives
 * each level's last column, 0-based (the reference's k_js minus one); k_0
 * (out) the number of zero columns.  Runs on the device (one workgroup). */
int dbg_adapt_low_rank_f64(double* M, int64_t lda, int32_t n, double* a, double* b, int32_t rank, int32_t* perm,
                           double eps1, int32_t* k_js, int32_t* k_0);

/* Covariance generators of the reference's simulation scripts
 * (scripts/simulations/simulations_functions.jl), built on the device:
 *   DBG_COV_FIXED_DENSE   fixed_dense(n; jitter)   :21-24   params = {jitter}
 *   DBG_COV_FIXED_SPARSE  fixed_sparse(n, n_sp)    :33-41   params = {n_sp}
 *   DBG_COV_MATERN        mattern_cov(n; nu, dist, sigma2)  :56-79
 *                         params = {nu, dist, sigma2}, nu in {0.5, 1.5, 2.5} (closed forms; the
 *                         reference's presets mattern_cov0/1/2 use nu = 1.5)
 *   DBG_COV_LOWRANK       B B' + diag(psi), B n x k column-major, psi n: the factor model, and
 *                         rand_spd(n; jitter) :13-18 with k = n, psi = jitter * n (A drawn by
 *                         the caller, whose RNG stream the library cannot reproduce)
 * standardize != 0 rescales to a correlation matrix (S_ij / (s_i s_j), ones on the
 * diagonal).  Sigma (n x n, lds) is device memory when sigma_on_device != 0 (it can
 * then go straight to dbg_plan_create), else host memory.  B and psi may be host or
 * device memory.  No reference counterpart as a library call. */
#define DBG_COV_FIXED_DENSE 0
#define DBG_COV_FIXED_SPARSE 1
#define DBG_COV_MATERN 2
#define DBG_COV_LOWRANK 3
int dbg_cov_build(int32_t kind, int32_t n, const double* params, int32_t k, const double* B, const double* psi,
                  int32_t standardize, double* Sigma, int64_t lds, int32_t sigma_on_device);

/* No counterpart in the reference: cholesky_genz! (as dbg_chol_genz_f64) of
 * a matrix already in device memory, for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops.  It allocates
 * nothing: M (n x n, lda), a, b, perm, scale and rank are the caller's device
 * buffers, factored in place on `stream` (a hipStream_t; NULL for the null
 * stream) without waiting, and the only scratch, the window's conditional
 * means and eps0, is a fixed array of the library (512 doubles, one per
 * device, held until the process ends).  Limits: one pivot window,
 * 1 <= n <= 511 and chol_block_size2 >= n (larger n or several windows:
 * DBG_ERR_ARG).  Calls that overlap on one device race on that array. */
int dbg_chol_genz_f64_dev_noalloc(double* M, int64_t lda, int32_t n, double* a, double* b, int32_t bs1,
                                  int32_t bs2, int32_t* perm, double* scale, int32_t* rank, void* stream);

/* cholesky_classic!(M, a, b, bs1, bs2, bs3, bs4)  src/cholesky.jl:144-242
 * Unpivoted, in place on host arrays, FP64 on the device. */
int dbg_chol_classic_f64(double* M, int64_t lda, int32_t n, double* a, double* b, int32_t bs1,
                         int32_t bs2, int32_t bs3, int32_t bs4, double* scale);

/* QMCData(C, a, b, opts, rng, qmc_type)        src/qmc.jl:149-202
 * QMCDataSparse(C, a, b, opts, rng, qmc_type)  src/qmc.jl:225-282
 * SparseQMC.QMCDataSparse(C, a, b; opts)       src/qmc_sparse.jl:58-90
 * Copies Sigma (n x n, lda; host or device memory), a and b (host) to the
 * device, factors (pivoted for
 * DBG_KIND_DENSE, unpivoted otherwise), applies truncate_matrixU! and builds
 * the packed / CSC layouts.  The random shifts are set separately (the RNG
 * stays with the caller).  Fails with DBG_ERR_ARG for NaN bounds (where the
 * reference would return a NaN estimate).
 * Where the reference throws for a rank-deficient factor (src/qmc.jl:195-197)
 * a dense plan instead takes the rank-deficient route (no reference
 * counterpart): cholesky_rowmax! (src/cholesky.jl:266-432) with the rank check
 * m_ii > n eps, the dependent columns' rows completed from Sigma,
 * adapt_low_rank! (src/cholesky.jl:20-123) with eps1 = sqrt(eps), and an FP64
 * integration over the rank r latent variables in which each latent
 * variable's interval is the intersection of the constraints of its level
 * (r - 1 sampled dimensions: the first r - 1 rows of the shifts are used).
 * dbg_plan_get_factor then returns the reordered factor, its columns' bounds
 * and variables and the rank; dbg_plan_lowrank_info the levels.  DBG_ERR_RANK
 * remains for a matrix of rank 0. */
int dbg_plan_create(dbg_plan** plan, const dbg_opts* opts, int32_t n, const double* Sigma,
                    int64_t lda, const double* a, const double* b);
int dbg_plan_destroy(dbg_plan* plan);

/* Redo the constructor's factorization (cholesky_genz! / cholesky_classic!,
 * truncate_matrixU!, packed and CSC layouts) from the device-resident copies
 * of Sigma, a and b: the factor part of QMCData (src/qmc.jl:189-199) for
 * callers that time or repeat the whole job with its inputs in HBM. */
int dbg_plan_refactor(dbg_plan* plan);

/* Random shifts: rand(rng, T, n-1, n_reps) (src/qmc_generators.jl:52) as
 * double, or rand(rng, UInt64, n-1, n_reps) (:63) for Sobol; column-major
 * (n-1) x n_reps.  For DBG_PREC_MIXED the Richtmyer shifts enter the point
 * map as 32-bit fixed-point fractions, floor(frac(shift) 2^32). */
int dbg_plan_set_shifts(dbg_plan* plan, const void* shifts);
/* Sobol direction integers, (n-1) x 32 row-major: V[d*32+k] = m_{k+1} << (31-k) */
int dbg_plan_set_sobol_directions(dbg_plan* plan, const uint32_t* V);

/* qmc_pnorm!(D)  src/qmc.jl:783-850 (SparseQMC.qmc_pnorm! src/qmc_sparse.jl:272-317
 * for DBG_KIND_SPARSEQMC).  Runs rounds of the integration until
 * n_pts >= max_pts or err3 < max_abs_err; the generator state carries over
 * between calls as in the reference. */
int dbg_qmc_pnorm(dbg_plan* plan, dbg_result* out);

/* One qmc_loop!(D, n_pts0, c_1, dc_1) (src/qmc.jl:315-444) over the current
 * block of points, queued on the plan's stream without waiting: the timed
 * hot path.  dbg_plan_sync waits for it. */
int dbg_qmc_loop(dbg_plan* plan, int64_t n_pts0);
int dbg_plan_sync(dbg_plan* plan);
/* device time of the integration kernel of the last dbg_qmc_loop (HIP events
 * on the plan's stream) */
int dbg_plan_last_kernel_ms(dbg_plan* plan, float* ms);
/* summary of the current per-shift means (after dbg_qmc_loop) */
int dbg_plan_summary(dbg_plan* plan, dbg_result* out);

/* Gradient of log P (no reference counterpart).  The estimate differentiated
 * is one round of the tail-form estimate over the plan's current block of m
 * points and its shifts (the forward is re-run with the per-sample y, the
 * conditional means and the weights recorded).  Outputs (any may be NULL):
 *   log_value  log P of that estimate;
 *   ga, gb     d log P / d a, d log P / d b of the inputs of dbg_plan_create
 *              (length n, original order);
 *   gSigma     d log P / d Sigma (n x n, column-major, symmetric): for a
 *              symmetric perturbation E, d log P = sum_ij gSigma_ij E_ij;
 *              the pivot order of the factorization is held fixed;
 *   ga_std, gb_std, gU  the same with respect to the plan's standardized,
 *              permuted bounds and its factor U (upper triangle, n x n).
 * Dense plans only.  The samples run in chunks: each chunk is recorded (y
 * and the conditional means, 2 (n-1) values of T per sample), swept back,
 * and folded into the sums with its share of the estimate, so the records
 * take at most the plan's gradient budget (dbg_plan_set_grad_budget) however
 * many points the round has.  dbg_plan_last_kernel_ms then reports the device
 * time of the whole gradient. */
int dbg_qmc_grad(dbg_plan* plan, double* log_value, double* ga, double* gb, double* gSigma,
                 double* ga_std, double* gb_std, double* gU);

/* The level structure of a dense plan (no reference counterpart): its rank r
 * (n for a plan of full rank), the last column of each level k_js (r entries,
 * 0-based: adapt_low_rank!'s k_js minus one; 0 .. n-1 at full rank) and the
 * number of zero-variance columns k_0 at the end.  Any output may be NULL.
 * The gradient (dbg_qmc_grad) is defined for plans of full rank only. */
int dbg_plan_lowrank_info(dbg_plan* plan, int32_t* rank, int32_t* k_js, int32_t* k_0);

/* Device memory for the gradient's y and s records (no reference
 * counterpart): record_bytes per chunk of samples (at least one tile of
 * samples per chunk), 0 for the default of 4 GiB.  Buffers of an earlier
 * gradient are released when the budget changes. */
int dbg_plan_set_grad_budget(dbg_plan* plan, int64_t record_bytes);
/* The chunking of the last gradient: samples per chunk (whole tiles), the
 * number of chunks and the device memory held by all of the gradient's
 * buffers.  Zeros before the first dbg_qmc_grad. */
int dbg_plan_grad_info(dbg_plan* plan, int64_t* chunk_samples, int32_t* n_chunks, int64_t* device_bytes);

/* The FP64 re-evaluation of the last round of a mixed dense plan (no
 * reference counterpart): tiles re-evaluated per shift, the largest share of
 * a shift's sum held by the tiles left in FP32 (the estimate differs from the
 * plain FP64 path's only through those tiles), and the device time of the
 * selection, the FP64 tiles and the merge (HIP events on the plan's stream).
 * Zeros for plans without it. */
int dbg_plan_refine_info(dbg_plan* plan, int32_t* tiles_per_shift, double* unrefined_share, float* ms);

/* The factor and the shift means held by the plan. */
int dbg_plan_get_factor(dbg_plan* plan, double* U, int32_t* perm, double* a, double* b,
                        double* scale, int32_t* rank);
int dbg_plan_get_reps(dbg_plan* plan, double* rep_mant, int32_t* rep_exp);
/* generator block state: Richtmyer n_base = state * m, Sobol first index =
 * skip + state * m; for DBG_KIND_SPARSEQMC the number of blocks generated */
int dbg_plan_get_state(dbg_plan* plan, int64_t* state);
int dbg_plan_set_state(dbg_plan* plan, int64_t state);

/* ---- batched problems of uneven dimension (no reference counterpart: the
 * reference calls qmc_pnorm!(QMCData(...)) once per problem) ----
 * K independent problems, each the QMCData(Sigma_k, a_k, b_k) constructor
 * plus qmc_pnorm! with max_pts = m (one round), Richtmyer points, erf form,
 * d_k <= 512 and inside one pivot window.  Inputs are concatenated:
 * sigmas = [Sigma_1 (d_1 x d_1, column-major), Sigma_2, ...], a and b =
 * [a_1, a_2, ...], shifts = [(d_1 - 1) x n_reps, ...].  dbg_batch_create
 * uploads, picks the dimension-block width with a timed trial on a sample of
 * the work (8, 16 or 32), and captures the whole job -- the batched device
 * factorizations, one persistent integration launch over all
 * (problem, shift, tile) work items largest-d first, and the per-problem
 * reductions -- into one HIP graph that dbg_batch_run replays.
 * results: K x 8 doubles per problem: value, log value, err3, rel_err3,
 * mean mantissa, err mantissa, binary exponent, rank (value NaN if the
 * factor is rank deficient). */
typedef struct dbg_batch dbg_batch;
int dbg_batch_create(dbg_batch** batch, const dbg_opts* opts, int32_t K, const int32_t* dims,
                     const double* sigmas, const double* a, const double* b, const double* shifts);
int dbg_batch_run(dbg_batch* batch, double* results);
/* device time of the last dbg_batch_run (HIP events around the graph) */
int dbg_batch_last_ms(dbg_batch* batch, float* ms);
/* chosen block width and the trial times of widths 8, 16, 32 (ms) */
int dbg_batch_tuning(dbg_batch* batch, int32_t* width, float* ms_w8_w16_w32);
/* Override of the autotuner's dimension-block width (8 or 16: vector-ALU
 * contraction; 32: for DBG_PREC_MIXED the FP16 matrix-core contraction): the
 * job's graph is captured again at that width (no reference counterpart) */
int dbg_batch_set_width(dbg_batch* batch, int32_t width);
int dbg_batch_destroy(dbg_batch* batch);

#ifdef __cplusplus
}
#endif
#endif /* DBGENZ_H */
