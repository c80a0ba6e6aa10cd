/*
 * genz_oracle.c -- CPU FP64 restatement of the Genz quasi-Monte Carlo
 * multivariate-normal probability path of JakobPeterlin/DBGenzPaper.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the pytest suite,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py call it to
 * produce the expected values the HIP library is compared against.  Nothing
 * in the shipped library links it, calls it or falls back to it.
 *
 * Every routine follows one reference routine (paths relative to the
 * reference repository root) and keeps its operation order where the
 * reference fixes one.  Where the reference delegates to a library whose
 * summation order is not fixed (BLAS mul!/ldiv!, Julia's pairwise sum,
 * LoopVectorization's @turbo), the oracle sums in ascending index order and
 * writes the reference's muladd as an explicit fma().  Everything else is
 * compiled with -ffp-contract=off.
 *
 * Products of per-dimension interval masses are carried as a mantissa in
 * [0.5, 1) and a binary exponent.  Scaling by a power of two is exact, so
 * whenever the reference's plain Float64 product does not underflow the
 * result is bit-identical to it; when it would underflow, the oracle still
 * returns log P.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Scaled non-negative-ish numbers: value = m * 2^e                           */
/* ------------------------------------------------------------------------ */
typedef struct {
    double m;
    int e;
} scaled_t;

static scaled_t sc_make(double v)
{
    scaled_t r;
    int e = 0;
    r.m = frexp(v, &e);
    r.e = (r.m == 0.0 || !isfinite(r.m)) ? 0 : e;
    if (!isfinite(v)) r.m = v;
    return r;
}

static double sc_to_double(scaled_t a) { return ldexp(a.m, a.e); }

/* common exponent of two scaled numbers (ignoring zero mantissas) */
static int sc_common_exp(scaled_t a, scaled_t b)
{
    if (a.m == 0.0) return b.e;
    if (b.m == 0.0) return a.e;
    return a.e > b.e ? a.e : b.e;
}

/* ------------------------------------------------------------------------ */
/* Special functions                                                          */
/* ------------------------------------------------------------------------ */

/* Wichura (1988), Algorithm AS 241 PPND16: normal quantile, ~1e-16 relative.
 * Used only as the starting point of the refined erfinv below. */
static double as241_ndtri(double p)
{
    const double q = p - 0.5;
    double r, val;
    if (fabs(q) <= 0.425) {
        r = 0.180625 - q * q;
        val = q * (((((((2.5090809287301226727e+3 * r + 3.3430575583588128105e+4) * r +
                        6.7265770927008700853e+4) * r + 4.5921953931549871457e+4) * r +
                      1.3731693765509461125e+4) * r + 1.9715909503065514427e+3) * r +
                    1.3314166789178437745e+2) * r + 3.3871328727963666080e+0) /
              (((((((5.2264952788528545610e+3 * r + 2.8729085735721942674e+4) * r +
                    3.9307895800092710610e+4) * r + 2.1213794301586595867e+4) * r +
                  5.3941960214247511077e+3) * r + 6.8718700749205790830e+2) * r +
                4.2313330701600911252e+1) * r + 1.0);
        return val;
    }
    r = (q < 0.0) ? p : 1.0 - p;
    if (r <= 0.0) return q < 0.0 ? -INFINITY : INFINITY;
    r = sqrt(-log(r));
    if (r <= 5.0) {
        r -= 1.6;
        val = (((((((7.74545014278341407640e-4 * r + 2.27238449892691845833e-2) * r +
                    2.41780725177450611770e-1) * r + 1.27045825245236838258e+0) * r +
                  3.64784832476320460504e+0) * r + 5.76949722146069140550e+0) * r +
                4.63033784615654529590e+0) * r + 1.42343711074968357734e+0) /
              (((((((1.05075007164441684324e-9 * r + 5.47593808499534494600e-4) * r +
                    1.51986665636164571966e-2) * r + 1.48103976427480074590e-1) * r +
                  6.89767334985100004550e-1) * r + 1.67638483018380384940e+0) * r +
                2.05319162663775882187e+0) * r + 1.0);
    } else {
        r -= 5.0;
        val = (((((((2.01033439929228813265e-7 * r + 2.71155556874348757815e-5) * r +
                    1.24266094738807843860e-3) * r + 2.65321895265761230930e-2) * r +
                  2.96560571828504891230e-1) * r + 1.78482653991729133580e+0) * r +
                5.46378491116411436990e+0) * r + 6.65790464350110377720e+0) /
              (((((((2.04426310338993978564e-15 * r + 1.42151175831644588870e-7) * r +
                    1.84631831751005468180e-5) * r + 7.86869131145613259100e-4) * r +
                  1.48753612908506148525e-2) * r + 1.36929880922735805310e-1) * r +
                5.99832206555887937690e-1) * r + 1.0);
    }
    return q < 0.0 ? -val : val;
}

/* erfinv(y): AS 241 start (about 1e-16 relative) and one Newton step on erf
 * (|y| <= 1/2) or on erfc (|y| > 1/2, where 1-|y| is exact), so the result
 * is within an ulp or two of the true inverse, the accuracy of
 * SpecialFunctions.erfinv that it stands in for (src/qmc.jl:386). */
double orc_erfinv(double y)
{
    if (isnan(y) || y < -1.0 || y > 1.0) return NAN;
    if (y == 1.0) return INFINITY;
    if (y == -1.0) return -INFINITY;
    if (y == 0.0) return y;
    const double ay = fabs(y);
    double x;
    if (ay <= 0.5)
        x = as241_ndtri(0.5 + 0.5 * ay) * 0.70710678118654752440;
    else
        x = -as241_ndtri(0.5 * (1.0 - ay)) * 0.70710678118654752440;
    const double k = 1.12837916709551257390; /* 2/sqrt(pi) */
    const double fp = k * exp(-x * x);
    if (fp > 0.0) {
        const double f = (ay <= 0.5) ? (erf(x) - ay) : ((1.0 - ay) - erfc(x));
        x -= f / fp;
    }
    return copysign(x, y);
}

double orc_as241(double p) { return as241_ndtri(p); }

/* ------------------------------------------------------------------------ */
/* Point generators  (src/qmc_generators.jl)                                 */
/* ------------------------------------------------------------------------ */

/* q_i = frac(sqrt(p_i)), p_1 = 2, p_2 = 3, ...   (get_sqrt_primes!, :300-310) */
int orc_sqrt_primes(int n, double* q)
{
    int found = 0;
    long p = 1;
    while (found < n) {
        /* nextprime(p + 1): smallest prime >= p + 1 */
        long c = p + 1;
        for (;; ++c) {
            int is_p = c >= 2;
            for (long f = 2; f * f <= c && is_p; ++f)
                if (c % f == 0) is_p = 0;
            if (is_p) break;
        }
        p = c;
        double s = sqrt((double)p);
        q[found++] = s - floor(s);
    }
    return 0;
}

/* X[j, i] = frac(Float64(n0 + j) * q_i), j 1-based  (richtmyer_mat!, :314-325) */
static inline double richtmyer_x(int64_t n0_plus_j, double q)
{
    double x = (double)n0_plus_j * q;
    return x - floor(x);
}

void orc_richtmyer_block(int n_dims, int64_t m, int64_t n0, const double* q, double* X)
{
    for (int i = 0; i < n_dims; ++i)
        for (int64_t j = 1; j <= m; ++j)
            X[(j - 1) + (size_t)i * m] = richtmyer_x(n0 + j, q[i]);
}

/* Sobol point with sequence index idx (the idx-th call of Sobol.jl next!):
 * XOR of the direction integers selected by the Gray code of idx.
 * V is n_dims x 32 (row per dimension), V[d*32 + k] = m_{k+1} << (31 - k).
 * Equals floor(x * 2^32) of copyt_as_UInt64! (src/qmc_generators.jl:340-352). */
static inline uint32_t sobol_x(const uint32_t* Vd, uint64_t idx)
{
    uint64_t g = idx ^ (idx >> 1);
    uint32_t x = 0;
    for (int k = 0; k < 32 && g; ++k, g >>= 1)
        if (g & 1u) x ^= Vd[k];
    return x;
}

void orc_sobol_block(int n_dims, const uint32_t* V, uint64_t first_idx, int64_t m, uint64_t* X)
{
    for (int i = 0; i < n_dims; ++i)
        for (int64_t j = 1; j <= m; ++j)
            X[(j - 1) + (size_t)i * m] = sobol_x(V + (size_t)i * 32, first_idx + (uint64_t)j);
}

/* ------------------------------------------------------------------------ */
/* Cholesky factorizations  (src/cholesky.jl)                                 */
/* ------------------------------------------------------------------------ */
#define MM(i, j) M[(size_t)(i) + (size_t)(j) * (size_t)n] /* 0-based, column-major */

/* ldiv!(UpperTriangular(M[r0:r1, r0:r1])', M[r0:r1, c0:c1]) then
 * mul!(M[c0:c1, c0:c1], transpose(M[r0:r1, c0:c1]), M[r0:r1, c0:c1], -1, 1)
 * (half-open ranges).  cholesky.jl:201-206, 589-596, 603-609, 617-623 */
static void trsm_gemm_update(double* M, int n, int r0, int r1, int c0, int c1)
{
    for (int c = c0; c < c1; ++c) {
        for (int r = r0; r < r1; ++r) {
            double acc = MM(r, c);
            for (int l = r0; l < r; ++l) acc -= MM(l, r) * MM(l, c);
            MM(r, c) = acc / MM(r, r);
        }
    }
    for (int c = c0; c < c1; ++c)
        for (int r = c0; r < c1; ++r) {
            double acc = MM(r, c);
            for (int l = r0; l < r1; ++l) acc -= MM(l, r) * MM(l, c);
            MM(r, c) = acc;
        }
}

/* Standardize to a correlation matrix, scale the bounds.  cholesky.jl:450-473
 * (mirror = 1 also writes the lower triangle, as cholesky_genz! does). */
static double standardize(double* M, int n, double* a, double* b, double* scale, int mirror)
{
    double maxd = 0.0;
    for (int j = 0; j < n; ++j) {
        double v = fabs(MM(j, j));
        if (v > maxd || isnan(v)) maxd = v;
    }
    const double eps0 = DBL_EPSILON * maxd;
    for (int j = 0; j < n; ++j) {
        double v = MM(j, j);
        scale[j] = v > 0 ? sqrt(v) : 1.0;
    }
    for (int j = 0; j < n; ++j) {
        const double s_j = scale[j];
        const double inv_s_j = 1.0 / s_j;
        MM(j, j) = 1.0;
        a[j] *= inv_s_j;
        b[j] *= inv_s_j;
        for (int i = 0; i < j; ++i) {
            double val = MM(i, j) * (inv_s_j / scale[i]);
            MM(i, j) = val;
            if (mirror) MM(j, i) = val;
        }
    }
    return eps0;
}

/* cholesky_genz!  src/cholesky.jl:445-628.  perm is 0-based.  gap_out (may be
 * NULL) receives, per pivot step, (second-best - best)/|best| of the order
 * vector, so a permutation mismatch can be traced to a near tie. */
int orc_cholesky_genz(double* M, int n, double* a, double* b, int bs1, int bs2, int bs3,
                      int bs4, int* perm, double* scale, int* rank_out, double* gap_out)
{
    const double c2 = 1.0 / sqrt(2.0);
    const double is2pi = 2.0 / sqrt(2.0 * M_PI);
    const double eps0 = standardize(M, n, a, b, scale, 1);
    double* d_vec = calloc((size_t)n, sizeof(double));
    double* s_vec = calloc((size_t)n, sizeof(double));
    double* order = calloc((size_t)n, sizeof(double));
    if (!d_vec || !s_vec || !order) return -1;
    int rank = n;
    for (int j = 0; j < n; ++j) perm[j] = j;
    if (gap_out)
        for (int j = 0; j < n; ++j) gap_out[j] = INFINITY;

    for (int iv0 = 0; iv0 < n && rank == n; iv0 += bs4) {
        const int iv1 = iv0 + bs4 < n ? iv0 + bs4 : n;
        for (int iii0 = iv0; iii0 < iv1 && rank == n; iii0 += bs3) {
            const int iii1 = iii0 + bs3 < iv1 ? iii0 + bs3 : iv1;
            for (int ii0 = iii0; ii0 < iii1 && rank == n; ii0 += bs2) {
                const int ii1 = ii0 + bs2 < iii1 ? ii0 + bs2 : iii1;
                for (int i0 = ii0; i0 < ii1 && rank == n; i0 += bs1) {
                    const int i1 = i0 + bs1 < ii1 ? i0 + bs1 : ii1;
                    for (int k = i0; k < ii1; ++k) d_vec[k] = 0.0;
                    for (int i = i0; i < i1; ++i) {
                        for (int k = i; k < ii1; ++k) {
                            double m_kk = MM(k, k) - d_vec[k];
                            m_kk = m_kk < 0 ? 0.0 : sqrt(m_kk);
                            const double a_k = (a[k] - s_vec[k]) / m_kk * c2;
                            const double b_k = (b[k] - s_vec[k]) / m_kk * c2;
                            order[k] = erf(b_k) - erf(a_k);
                        }
                        /* argmin: first minimum; NaN counts as smaller than everything */
                        int i_m = i;
                        for (int k = i + 1; k < ii1; ++k) {
                            if (isnan(order[i_m])) break;
                            if (isnan(order[k]) || order[k] < order[i_m]) i_m = k;
                        }
                        if (gap_out) {
                            double second = INFINITY;
                            for (int k = i; k < ii1; ++k)
                                if (k != i_m && order[k] < second) second = order[k];
                            gap_out[i] = (second - order[i_m]) / fabs(order[i_m]);
                        }
                        const double d_ii = order[i_m];
                        double m_ii = MM(i_m, i_m) - d_vec[i_m];
                        if (i_m != i) {
                            int tp = perm[i]; perm[i] = perm[i_m]; perm[i_m] = tp;
                            double t;
                            t = d_vec[i]; d_vec[i] = d_vec[i_m]; d_vec[i_m] = t;
                            t = s_vec[i]; s_vec[i] = s_vec[i_m]; s_vec[i_m] = t;
                            t = a[i]; a[i] = a[i_m]; a[i_m] = t;
                            t = b[i]; b[i] = b[i_m]; b[i_m] = t;
                            for (int k = i0; k < n; ++k) {
                                t = MM(i, k); MM(i, k) = MM(i_m, k); MM(i_m, k) = t;
                            }
                            for (int k = 0; k < n; ++k) {
                                t = MM(k, i); MM(k, i) = MM(k, i_m); MM(k, i_m) = t;
                            }
                        }
                        if (m_ii > eps0) {
                            m_ii = sqrt(m_ii);
                        } else {
                            rank = i;
                            break;
                        }
                        MM(i, i) = m_ii;
                        const double a_norm = (a[i] - s_vec[i]) / m_ii;
                        const double b_norm = (b[i] - s_vec[i]) / m_ii;
                        const double y_i =
                            (exp(-(a_norm * a_norm) / 2) - exp(-(b_norm * b_norm) / 2)) * is2pi / d_ii;
                        if (i == i0) {
                            for (int j = i + 1; j < ii1; ++j) {
                                const double m_ij = MM(i, j) / m_ii;
                                MM(i, j) = m_ij;
                                MM(j, i) = m_ij;
                                d_vec[j] += m_ij * m_ij;
                                s_vec[j] += y_i * m_ij;
                            }
                        } else if (i < ii1 - 1) {
                            for (int r = i + 1; r < ii1; ++r) {
                                double acc = MM(r, i);
                                for (int l = i0; l < i; ++l) acc -= MM(r, l) * MM(l, i);
                                MM(r, i) = acc;
                            }
                            for (int j = i + 1; j < ii1; ++j) {
                                const double m_ij = MM(j, i) / m_ii;
                                MM(i, j) = m_ij;
                                MM(j, i) = m_ij;
                                d_vec[j] += m_ij * m_ij;
                                s_vec[j] += y_i * m_ij;
                            }
                        }
                    }
                    if (rank < n) break;
                    if (i1 < ii1) {
                        for (int c = i1; c < ii1; ++c)
                            for (int r = i1; r < ii1; ++r) {
                                double acc = MM(r, c);
                                for (int l = i0; l < i1; ++l) acc -= MM(r, l) * MM(l, c);
                                MM(r, c) = acc;
                            }
                    }
                }
                if (rank < n) break;
                if (ii1 < iii1) trsm_gemm_update(M, n, ii0, ii1, ii1, iii1);
            }
            if (rank < n) break;
            if (iii1 < iv1) trsm_gemm_update(M, n, iii0, iii1, iii1, iv1);
        }
        if (rank < n) break;
        if (iv1 < n) trsm_gemm_update(M, n, iv0, iv1, iv1, n);
    }
    free(d_vec);
    free(s_vec);
    free(order);
    *rank_out = rank;
    return 0;
}

/* cholesky_rowmax!  src/cholesky.jl:266-432.  The loop nest of cholesky_genz!
 * with diagonal pivoting: at step i the pivot is the largest m_kk = M[k,k] -
 * d[k] over k in [i, ii1), scanned from i with a strict > (:319-327, so the
 * first maximum wins and a NaN after i is never taken); the rank check
 * m_ii > 0 (no eps0) comes before the swap (:329-336); a and b are
 * standardized (:285-300) but not permuted.  perm is 0-based.  tol: the
 * rank check is m_ii > tol (0 for the reference; the rank-deficient route
 * of the plans uses n * DBL_EPSILON, LAPACK dpstrf's default). */
int orc_cholesky_rowmax(double* M, int n, double* a, double* b, int bs1, int bs2, int bs3,
                        int bs4, int* perm, double* scale, int* rank_out, double tol)
{
    (void)standardize(M, n, a, b, scale, 1);
    double* d_vec = calloc((size_t)n, sizeof(double));
    if (!d_vec) return -1;
    int rank = n;
    for (int j = 0; j < n; ++j) perm[j] = j;

    for (int iv0 = 0; iv0 < n && rank == n; iv0 += bs4) {
        const int iv1 = iv0 + bs4 < n ? iv0 + bs4 : n;
        for (int iii0 = iv0; iii0 < iv1 && rank == n; iii0 += bs3) {
            const int iii1 = iii0 + bs3 < iv1 ? iii0 + bs3 : iv1;
            for (int ii0 = iii0; ii0 < iii1 && rank == n; ii0 += bs2) {
                const int ii1 = ii0 + bs2 < iii1 ? ii0 + bs2 : iii1;
                for (int i0 = ii0; i0 < ii1 && rank == n; i0 += bs1) {
                    const int i1 = i0 + bs1 < ii1 ? i0 + bs1 : ii1;
                    for (int k = i0; k < ii1; ++k) d_vec[k] = 0.0;
                    for (int i = i0; i < i1; ++i) {
                        int i_m = i;
                        double m_ii = MM(i, i) - d_vec[i];
                        for (int k = i; k < ii1; ++k) {
                            const double m_kk = MM(k, k) - d_vec[k];
                            if (m_kk > m_ii) {
                                m_ii = m_kk;
                                i_m = k;
                            }
                        }
                        if (m_ii > tol) {
                            m_ii = sqrt(m_ii);
                        } else {
                            rank = i;
                            break;
                        }
                        if (i_m != i) {
                            int tp = perm[i]; perm[i] = perm[i_m]; perm[i_m] = tp;
                            double t;
                            t = d_vec[i]; d_vec[i] = d_vec[i_m]; d_vec[i_m] = t;
                            for (int k = i0; k < n; ++k) {
                                t = MM(i, k); MM(i, k) = MM(i_m, k); MM(i_m, k) = t;
                            }
                            for (int k = 0; k < n; ++k) {
                                t = MM(k, i); MM(k, i) = MM(k, i_m); MM(k, i_m) = t;
                            }
                        }
                        MM(i, i) = m_ii;
                        if (i == i0) {
                            for (int j = i + 1; j < ii1; ++j) {
                                const double m_ij = MM(i, j) / m_ii;
                                MM(i, j) = m_ij;
                                MM(j, i) = m_ij;
                                d_vec[j] += m_ij * m_ij;
                            }
                        } else if (i < ii1 - 1) {
                            for (int r = i + 1; r < ii1; ++r) {
                                double acc = MM(r, i);
                                for (int l = i0; l < i; ++l) acc -= MM(r, l) * MM(l, i);
                                MM(r, i) = acc;
                            }
                            for (int j = i + 1; j < ii1; ++j) {
                                const double m_ij = MM(j, i) / m_ii;
                                MM(i, j) = m_ij;
                                MM(j, i) = m_ij;
                                d_vec[j] += m_ij * m_ij;
                            }
                        }
                    }
                    if (rank < n) break;
                    if (i1 < ii1) {
                        for (int c = i1; c < ii1; ++c)
                            for (int r = i1; r < ii1; ++r) {
                                double acc = MM(r, c);
                                for (int l = i0; l < i1; ++l) acc -= MM(r, l) * MM(l, c);
                                MM(r, c) = acc;
                            }
                    }
                }
                if (rank < n) break;
                if (ii1 < iii1) trsm_gemm_update(M, n, ii0, ii1, ii1, iii1);
            }
            if (rank < n) break;
            if (iii1 < iv1) trsm_gemm_update(M, n, iii0, iii1, iii1, iv1);
        }
        if (rank < n) break;
        if (iv1 < n) trsm_gemm_update(M, n, iv0, iv1, iv1, n);
    }
    free(d_vec);
    *rank_out = rank;
    return 0;
}

/* adapt_low_rank!  src/cholesky.jl:20-123, sequentially as written (perm
 * 0-based; k_js receives each level's last column, 0-based). */
int orc_adapt_low_rank(double* M, int n, double* a, double* b, int rank, int* perm, double eps1,
                       int* k_js, int* k0_out)
{
    double* tmp = malloc((size_t)n * sizeof(double));
    if (!tmp) return -1;
    for (int i = 0; i < rank; ++i) k_js[i] = i;
    int k0 = 0, j = rank, n_active = n;
    while (j < n_active) {
        int ij = -1;
        for (int i = rank - 1; i >= 0; --i)
            if (fabs(MM(i, j)) > eps1) {
                ij = i;
                break;
            }
        if (ij < 0) {
            const int last = n_active - 1;
            if (j != last) {
                for (int r = 0; r < n; ++r) {
                    const double t = MM(r, last); MM(r, last) = MM(r, j); MM(r, j) = t;
                }
                double t = a[j]; a[j] = a[last]; a[last] = t;
                t = b[j]; b[j] = b[last]; b[last] = t;
                const int p = perm[j]; perm[j] = perm[last]; perm[last] = p;
            }
            k0 += 1;
            n_active -= 1;
            continue;
        }
        const double m_ij = MM(ij, j);
        const double inv_m = 1.0 / m_ij;
        for (int k = 0; k <= ij; ++k) MM(k, j) *= inv_m;
        a[j] *= inv_m;
        b[j] *= inv_m;
        if (m_ij < 0.0) {
            const double t = a[j]; a[j] = b[j]; b[j] = t;
        }
        const int pos = k_js[ij] + 1;
        if (j != pos) {
            for (int r = 0; r < n; ++r) tmp[r] = MM(r, j);
            const double va = a[j], vb = b[j];
            const int vp = perm[j];
            for (int c = j; c > pos; --c) {
                for (int r = 0; r < n; ++r) MM(r, c) = MM(r, c - 1);
                a[c] = a[c - 1];
                b[c] = b[c - 1];
                perm[c] = perm[c - 1];
            }
            for (int r = 0; r < n; ++r) MM(r, pos) = tmp[r];
            a[pos] = va;
            b[pos] = vb;
            perm[pos] = vp;
        }
        for (int k = ij; k < rank; ++k) k_js[k] += 1;
        j += 1;
    }
    free(tmp);
    *k0_out = k0;
    return 0;
}

/* cholesky_classic!  src/cholesky.jl:144-242 (no pivoting, no rank check). */
int orc_cholesky_classic(double* M, int n, double* a, double* b, int bs1, int bs2, int bs3,
                         int bs4, double* scale)
{
    (void)standardize(M, n, a, b, scale, 0);
    for (int iv0 = 0; iv0 < n; iv0 += bs4) {
        const int iv1 = iv0 + bs4 < n ? iv0 + bs4 : n;
        for (int iii0 = iv0; iii0 < iv1; iii0 += bs3) {
            const int iii1 = iii0 + bs3 < iv1 ? iii0 + bs3 : iv1;
            for (int ii0 = iii0; ii0 < iii1; ii0 += bs2) {
                const int ii1 = ii0 + bs2 < iii1 ? ii0 + bs2 : iii1;
                for (int i0 = ii0; i0 < ii1; i0 += bs1) {
                    const int i1 = i0 + bs1 < ii1 ? i0 + bs1 : ii1;
                    for (int i = i0; i < i1; ++i) {
                        double m_ii = MM(i, i);
                        if (i > i0) {
                            double dt = 0.0;
                            for (int l = i0; l < i; ++l) dt += MM(l, i) * MM(l, i);
                            m_ii -= dt;
                        }
                        m_ii = sqrt(m_ii);
                        MM(i, i) = m_ii;
                        if (i > i0 && i < i1 - 1) {
                            for (int j = i + 1; j < i1; ++j) {
                                double c = 0.0;
                                for (int l = i0; l < i; ++l) c += MM(l, j) * MM(l, i);
                                MM(i, j) -= c;
                                MM(j, i) = 0.0;
                            }
                        }
                        for (int j = i + 1; j < i1; ++j) MM(i, j) /= m_ii;
                    }
                    if (i1 < ii1) trsm_gemm_update(M, n, i0, i1, i1, ii1);
                }
                if (ii1 < iii1) trsm_gemm_update(M, n, ii0, ii1, ii1, iii1);
            }
            if (iii1 < iv1) trsm_gemm_update(M, n, iii0, iii1, iii1, iv1);
        }
        if (iv1 < n) trsm_gemm_update(M, n, iv0, iv1, iv1, n);
    }
    return 0;
}

/* truncate_matrixU!  src/qmc.jl:133-142 */
void orc_truncate_U(double* U, int n, double eps_T)
{
    const double min_val = eps_T / (double)(n * 9);
    for (size_t k = 0; k < (size_t)n * n; ++k)
        if (fabs(U[k]) < min_val) U[k] = 0.0;
}
#undef MM

/* ------------------------------------------------------------------------ */
/* Integration loops  (src/qmc.jl, src/qmc_sparse.jl)                         */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n;            /* dimension d */
    int64_t m;        /* points per shift and round */
    int n_reps;       /* number of random shifts */
    int gen;          /* 0 Richtmyer lattice, 1 Sobol */
    int phi_form;     /* 0 erf form (QMCData), 1 erfc form (QMCData_Distributions), 2 tail form */
    int bs_i, bs_i2, bs_j;
    int n_bits;       /* Sobol bits L */
    int threads;      /* OpenMP threads, <= 0: all */
} orc_desc;

/* Point source of one round (block) of m points. */
typedef struct {
    const double* q;        /* Richtmyer: (n-1) sqrt-prime fractions */
    const double* shifts_f; /* Richtmyer: (n-1) x n_reps, column-major */
    int64_t n_base;         /* Richtmyer: point j uses n_base + j */
    const uint32_t* V;      /* Sobol: (n-1) x 32 direction integers */
    const uint64_t* shifts_u; /* Sobol: (n-1) x n_reps */
    uint64_t sobol_first;   /* Sobol: point j uses sequence index sobol_first + j */
} orc_points;

static const double EPS_T = DBL_EPSILON;

/* The shifted uniform of one (point, dimension, shift) before the tent map:
 * shift mod 1 (Richtmyer, src/qmc_generators.jl:158-159) or digital shift
 * (Sobol, :197-198). */
static inline double point_u(const orc_desc* D, const orc_points* P, int i /*1-based*/,
                             int k /*0-based shift*/, int64_t j /*1-based point*/)
{
    const int nd = D->n - 1;
    double u;
    if (D->gen == 0) {
        const double shift_i = P->shifts_f[(size_t)(i - 1) + (size_t)k * nd];
        u = shift_i + richtmyer_x(P->n_base + j, P->q[i - 1]);
        u = u > 1.0 ? u - 1.0 : u;
    } else {
        const int L = D->n_bits;
        const uint64_t mask = (L >= 64) ? ~0ull : ((1ull << L) - 1);
        const double invscale = ldexp(1.0, -L);
        const uint64_t shift_i = P->shifts_u[(size_t)(i - 1) + (size_t)k * nd] & mask;
        const uint64_t x_ij = (uint64_t)sobol_x(P->V + (size_t)(i - 1) * 32, P->sobol_first + (uint64_t)j) & mask;
        u = (double)((x_ij ^ shift_i) & mask) * invscale;
    }
    return u;
}

/* Lower-tail normal quantile Phi^{-1}(t) for 0 < t <= 1/2: AS 241 start and one
 * Newton step on Phi(x) = erfc(-x / sqrt 2) / 2 (relative, no cancellation). */
double orc_ndtri_tail(double t)
{
    double x = as241_ndtri(t);
    const double phi = exp(-0.5 * x * x) * 0.39894228040143267794;
    if (phi > 0.0 && isfinite(x)) {
        const double f = 0.5 * erfc(-x * 0.70710678118654752440) - t;
        x -= f / phi;
    }
    return x;
}

/* Tail form (phi_form 2; no reference counterpart): the interval of the
 * dimension is kept as the normal tail masses ea = erfc(|za|)/2,
 * eb = erfc(|zb|)/2 beyond the erf arguments za, zb (no [-9, 9] clamp):
 *   zb <= 0:  Phi(a') = ea, dc = eb - ea        (mode 0)
 *   za >= 0:  1 - Phi(b') = eb, dc = ea - eb    (mode 1)
 *   else:     dc = 1 - ea - eb                  (mode 2)
 * and the sample's tail mass comes from the nearer tail. */
static inline double tail_sample(double w, double ea, double eb, double dc, int mode)
{
    const double u_lo = fma(w, dc, ea);
    const double u_hi = fma(1.0 - w, dc, eb);
    const int upper = (mode == 1) || (mode == 2 && u_hi < u_lo);
    const double t = fmax(upper ? u_hi : u_lo, 1e-300);
    const double y = orc_ndtri_tail(t);
    return upper ? -y : y;
}

static inline void tail_masses(double za, double zb, double* ea, double* eb, double* dc, int* mode)
{
    *ea = 0.5 * erfc(fabs(za));
    *eb = 0.5 * erfc(fabs(zb));
    if (zb <= 0.0) {
        *mode = 0;
        *dc = *eb - *ea;
    } else if (za >= 0.0) {
        *mode = 1;
        *dc = *ea - *eb;
    } else {
        *mode = 2;
        *dc = (1.0 - *ea) - *eb;
    }
}

/* rand_points!: shift mod 1 (Richtmyer) or digital shift (Sobol), tent,
 * affine map to [c, c+dc], clamp, 2u-1.  src/qmc_generators.jl:137-204 */
static inline double rand_point(const orc_desc* D, const orc_points* P, int i /*1-based*/,
                                int k /*0-based shift*/, int64_t j /*1-based point*/, double c,
                                double dc)
{
    const int nd = D->n - 1;
    double u;
    if (D->gen == 0) {
        const double shift_i = P->shifts_f[(size_t)(i - 1) + (size_t)k * nd];
        u = shift_i + richtmyer_x(P->n_base + j, P->q[i - 1]);
        u = u > 1.0 ? u - 1.0 : u;
    } else {
        const int L = D->n_bits;
        const uint64_t mask = (L >= 64) ? ~0ull : ((1ull << L) - 1);
        const double invscale = ldexp(1.0, -L);
        const uint64_t shift_i = P->shifts_u[(size_t)(i - 1) + (size_t)k * nd] & mask;
        const uint64_t x_ij = (uint64_t)sobol_x(P->V + (size_t)(i - 1) * 32, P->sobol_first + (uint64_t)j) & mask;
        u = (double)((x_ij ^ shift_i) & mask) * invscale;
    }
    u = fabs(fma(u, 2.0, -1.0));
    u = fma(u, dc, c);
    u = fmin(fmax(u, EPS_T), 1.0 - EPS_T);
    return fma(2.0, u, -1.0);
}

/* c, dc of the next dimension from the erf arguments of its interval,
 * c_val = (a - s) / (U_ii sqrt 2) and dc_val = (b - s) / (U_ii sqrt 2) for the
 * conditional mean s.  src/qmc.jl:405-429 (erf form) and :579-593 (erfc form). */
static inline void phi_z_update(int phi_form, double c_val, double dc_val, double* c_out, double* dc_out)
{
    c_val = fmax(fmin(c_val, 9.0), -9.0);
    dc_val = fmax(fmin(dc_val, 9.0), -9.0);
    if (phi_form == 0) {
        c_val = erf(c_val);
        dc_val = erf(dc_val);
        c_val = 0.5 * (1.0 + c_val);
        dc_val = 0.5 * (1.0 + dc_val) - c_val;
    } else {
        c_val = 0.5 * erfc(-c_val);
        dc_val = 0.5 * erfc(-dc_val) - c_val;
    }
    *c_out = c_val;
    *dc_out = dc_val;
}

/* y of one dimension for either form (state c/ea, dc, eb, mode). */
static inline double draw_y(const orc_desc* D, const orc_points* P, int i, int k, int64_t j, double c,
                            double dc, double eb, int mode)
{
    if (D->phi_form == 2) {
        const double w = fabs(fma(point_u(D, P, i, k, j), 2.0, -1.0));
        return tail_sample(w, c, eb, dc, mode);
    }
    return sqrt(2.0) * orc_erfinv(rand_point(D, P, i, k, j, c, dc));
}

/* the state of an interval given by its erf arguments za <= zb */
static inline void phi_state_z(int phi_form, double za, double zb, double* c_out, double* dc_out,
                               double* eb_out, int* mode_out)
{
    if (phi_form == 2)
        tail_masses(za, zb, c_out, eb_out, dc_out, mode_out);
    else
        phi_z_update(phi_form, za, zb, c_out, dc_out);
}

static inline void phi_state(int phi_form, double s_val, double a_i, double b_i, double inv,
                             double* c_out, double* dc_out, double* eb_out, int* mode_out)
{
    phi_state_z(phi_form, (a_i - s_val) * inv, (b_i - s_val) * inv, c_out, dc_out, eb_out, mode_out);
}

/* Sum of p_t * 2^{e_t} over a batch, as a scaled number. */
static scaled_t batch_sum(const double* pm, const int* pe, int len)
{
    int E = INT32_MIN;
    for (int t = 0; t < len; ++t)
        if (pm[t] != 0.0 && pe[t] > E) E = pe[t];
    scaled_t r = {0.0, 0};
    if (E == INT32_MIN) {
        double s = 0.0; /* all zero (or non-finite): plain sum keeps NaN */
        for (int t = 0; t < len; ++t) s += pm[t];
        r.m = s;
        return r;
    }
    double s = 0.0;
    for (int t = 0; t < len; ++t) s += ldexp(pm[t], pe[t] - E);
    r.m = s;
    r.e = E;
    return r;
}

static scaled_t sc_add(scaled_t a, scaled_t b)
{
    const int E = sc_common_exp(a, b);
    scaled_t r;
    r.m = ldexp(a.m, a.e - E) + ldexp(b.m, b.e - E);
    r.e = E;
    return r;
}

/* running-mean merge  src/qmc.jl:437-442 :
 * qmc_reps[k] = muladd(qmc_reps[k], n0/(n0+m), sum_p * inv(n0+m)) */
static scaled_t rep_merge(scaled_t rep, scaled_t sum_p, int64_t n_pts0, int64_t m)
{
    const int64_t n_pts_total = n_pts0 + m;
    const double inv_total = 1.0 / (double)n_pts_total;
    const double old_weight = (double)n_pts0 * inv_total;
    scaled_t mean = sum_p;
    mean.m = mean.m * inv_total;
    const int E = sc_common_exp(rep, mean);
    scaled_t r;
    r.m = fma(ldexp(rep.m, rep.e - E), old_weight, ldexp(mean.m, mean.e - E));
    r.e = E;
    return r;
}

/* Dense qmc_loop!  src/qmc.jl:315-444 (phi_form 0) and :485-608 (phi_form 1).
 * U: n x n column-major (only the strict upper part and the diagonal are read);
 * a, b: permuted standardized bounds; rep_m/rep_e: per-shift running means. */
int orc_qmc_loop_dense(const orc_desc* D, const orc_points* P, const double* U, const double* a,
                       const double* b, int64_t n_pts0, double c_1, double dc_1, double eb_1,
                       int mode_1, double* rep_m, int* rep_e)
{
    const int n = D->n;
    const int64_t m = D->m;
    const int bs_j = D->bs_j, bs_i = D->bs_i, bs_i2 = D->bs_i2;
    const double c2 = sqrt(2.0);
    const int64_t n_batches = (m + bs_j - 1) / bs_j;
    scaled_t* bsum = malloc(sizeof(scaled_t) * (size_t)n_batches);
    if (!bsum) return -1;
#define U1(r, c) U[(size_t)((r)-1) + (size_t)((c)-1) * (size_t)n]
    for (int k = 0; k < D->n_reps; ++k) {
#pragma omp parallel num_threads(D->threads > 0 ? D->threads : omp_get_max_threads())
        {
            double* Y = malloc(sizeof(double) * (size_t)bs_j * (size_t)(n > 1 ? n - 1 : 1));
            double* s = calloc((size_t)bs_j * (size_t)(bs_i > 0 ? bs_i : 1), sizeof(double));
            double* cv = malloc(sizeof(double) * (size_t)bs_j);
            double* dcv = malloc(sizeof(double) * (size_t)bs_j);
            double* ebv = malloc(sizeof(double) * (size_t)bs_j);
            int* mdv = malloc(sizeof(int) * (size_t)bs_j);
            double* pm = malloc(sizeof(double) * (size_t)bs_j);
            int* pe = malloc(sizeof(int) * (size_t)bs_j);
#define Y1(t, i) Y[(size_t)((t)-1) + (size_t)((i)-1) * (size_t)bs_j]
#define S1(t, c) s[(size_t)((t)-1) + (size_t)((c)-1) * (size_t)bs_j]
#pragma omp for schedule(static)
            for (int64_t bi = 0; bi < n_batches; ++bi) {
                const int64_t j1 = bi * bs_j + 1;
                const int64_t j2 = j1 + bs_j - 1 < m ? j1 + bs_j - 1 : m;
                const int len = (int)(j2 - j1 + 1);
                scaled_t p0 = sc_make(dc_1);
                for (int t = 0; t < bs_j; ++t) {
                    cv[t] = c_1;
                    dcv[t] = dc_1;
                    ebv[t] = eb_1;
                    mdv[t] = mode_1;
                    pm[t] = p0.m;
                    pe[t] = p0.e;
                }
                memset(s, 0, sizeof(double) * (size_t)bs_j * (size_t)(bs_i > 0 ? bs_i : 1));
                for (int ii1 = 1; ii1 <= n - 1; ii1 += bs_i) {
                    const int ii2 = ii1 + bs_i - 1 < n - 1 ? ii1 + bs_i - 1 : n - 1;
                    if (ii1 > 1) {
                        /* s[:, 1:ncols] = Y[:, 1:ii1-1] * U[1:ii1-1, ii1+1:ii2+1] */
                        /* (per element: acc = 0, then fma over l ascending; the loop
                         * nest keeps t innermost so it vectorizes) */
                        const int ncols = ii2 - ii1 + 1;
                        for (int cc = 1; cc <= ncols; ++cc)
                            for (int t = 1; t <= len; ++t) S1(t, cc) = 0.0;
                        for (int cc0 = 1; cc0 <= ncols; cc0 += 16) {
                            const int cc1 = cc0 + 15 < ncols ? cc0 + 15 : ncols;
                            for (int l = 1; l <= ii1 - 1; ++l)
                                for (int cc = cc0; cc <= cc1; ++cc) {
                                    const double u = U1(l, ii1 + cc);
                                    for (int t = 1; t <= len; ++t) S1(t, cc) = fma(Y1(t, l), u, S1(t, cc));
                                }
                        }
                    }
                    for (int i1 = ii1; i1 <= ii2; i1 += bs_i2) {
                        const int i2 = i1 + bs_i2 - 1 < ii2 ? i1 + bs_i2 - 1 : ii2;
                        if (i1 > ii1) {
                            const int prev1 = ii1 > i1 - bs_i2 ? ii1 : i1 - bs_i2;
                            const int prev2 = i1 - 1;
                            for (int cc = i1 - ii1 + 1; cc <= ii2 - ii1 + 1; ++cc)
                                for (int l = prev1; l <= prev2; ++l) {
                                    const double u = U1(l, ii1 + cc);
                                    for (int t = 1; t <= len; ++t) S1(t, cc) = fma(Y1(t, l), u, S1(t, cc));
                                }
                        }
                        for (int i = i1; i <= i2; ++i) {
                            for (int t = 1; t <= len; ++t) {
                                Y1(t, i) = draw_y(D, P, i, k, j1 + t - 1, cv[t - 1], dcv[t - 1], ebv[t - 1],
                                                  mdv[t - 1]);
                            }
                            const int col = i - ii1 + 1;
                            const int K = i - i1 + 1;
                            for (int kk = 1; kk <= K; ++kk) {
                                const double u = U1(i1 + kk - 1, i + 1);
                                for (int t = 1; t <= len; ++t) S1(t, col) = fma(Y1(t, i1 + kk - 1), u, S1(t, col));
                            }
                            const double a_i = a[i], b_i = b[i]; /* a[i+1], 1-based */
                            const double inv = 1.0 / (U1(i + 1, i + 1) * c2);
                            for (int t = 1; t <= len; ++t) {
                                phi_state(D->phi_form, S1(t, col), a_i, b_i, inv, &cv[t - 1], &dcv[t - 1],
                                          &ebv[t - 1], &mdv[t - 1]);
                                const double dcval = dcv[t - 1];
                                int e;
                                pm[t - 1] = frexp(pm[t - 1] * dcval, &e);
                                pe[t - 1] += e;
                            }
                        }
                    }
                }
                bsum[bi] = batch_sum(pm, pe, len);
            }
#undef Y1
#undef S1
            free(Y); free(s); free(cv); free(dcv); free(ebv); free(mdv); free(pm); free(pe);
        }
        scaled_t tot = {0.0, 0};
        for (int64_t bi = 0; bi < n_batches; ++bi) tot = sc_add(tot, bsum[bi]);
        scaled_t rep = {rep_m[k], rep_e[k]};
        rep = rep_merge(rep, tot, n_pts0, m);
        rep_m[k] = rep.m;
        rep_e[k] = rep.e;
    }
#undef U1
    free(bsum);
    return 0;
}

/* The first dimension's state, analytic (src/qmc.jl:800-809): erf arguments
 * arg_a = a_1 / (U_11 sqrt 2), arg_b likewise, unclamped. */
static void first_state(int phi_form, int kind, double arg_a, double arg_b, double* c_1, double* dc_1,
                        double* eb_1, int* mode_1)
{
    if (phi_form == 2 && kind != 2) {
        /* tail form: masses of the two tails */
        tail_masses(arg_a, arg_b, c_1, eb_1, dc_1, mode_1);
        if (*mode_1 == 2) *dc_1 = 0.5 * (erf(arg_b) - erf(arg_a));
    } else if (phi_form == 1 && kind == 0) {
        *c_1 = 0.5 * erfc(-arg_a);
        *dc_1 = 0.5 * (erfc(-arg_b) - erfc(-arg_a));
    } else {
        *c_1 = 0.5 * (1.0 + erf(arg_a));
        *dc_1 = 0.5 * (erf(arg_b) - erf(arg_a));
    }
}

/* ------------------------------------------------------------------------ */
/* Rank-deficient Sigma (no reference counterpart: QMCData throws for it,   */
/* src/qmc.jl:195-197)                                                      */
/* ------------------------------------------------------------------------ */
/* After cholesky_rowmax! with a rank tolerance (rank r), the rows 0 .. r-1
 * of the dependent columns completed from Sigma, and adapt_low_rank!
 * (src/cholesky.jl:20-123), column block i = beg[i] .. beg[i+1]-1 (i < r)
 * holds latent variable i in its first column (U[i, beg[i]] = U_ii) and the
 * dependent variables whose last coefficient above row r is on latent i
 * (scaled so that it is about 1, bounds exchanged where it was negative);
 * columns beg[r] .. n-1 have no coefficient (zero variance).  One sample's
 * integrand is prod_i dc_i, the interval of latent i being the intersection
 * over its block of
 *     [(a_c - s_c) inv_c, (b_c - s_c) inv_c],  inv_c = 1 / (U[i, c] sqrt 2),
 *     s_c = sum_{l < i} U[l, c] y_l            (fma, l ascending),
 * in erf-argument units (an empty intersection gives dc_i = 0), and y_g is
 * drawn from latent g's interval with point dimension g + 1 (1-based), as in
 * the dense loop (which this is when every block has one column).  The zero-
 * variance columns contribute the factor [a_c <= 0 <= b_c] to latent 0's
 * analytic interval (lowrank_first). */
static void lowrank_first(int phi_form, const double* U, int n, const double* a, const double* b,
                          const int* beg, int r, double* c_1, double* dc_1, double* eb_1, int* mode_1)
{
    const double sq2 = sqrt(2.0);
    double za = -INFINITY, zb = INFINITY;
    for (int c = beg[0]; c < beg[1]; ++c) {
        const double u = U[(size_t)c * n]; /* row 0 */
        za = fmax(za, a[c] / (u * sq2));
        zb = fmin(zb, b[c] / (u * sq2));
    }
    if (!(za <= zb)) zb = za;
    *eb_1 = 0.0;
    *mode_1 = 0;
    first_state(phi_form, 0, za, zb, c_1, dc_1, eb_1, mode_1);
    for (int c = beg[r]; c < n; ++c)
        if (!(a[c] <= 0.0 && 0.0 <= b[c])) *dc_1 = 0.0;
}

int orc_qmc_loop_lowrank(const orc_desc* D, const orc_points* P, const double* U, const double* a,
                         const double* b, const double* inv, const int* beg, int r, int64_t n_pts0,
                         double c_1, double dc_1, double eb_1, int mode_1, double* rep_m, int* rep_e)
{
    const int n = D->n;
    const int64_t m = D->m;
    const int bs_j = D->bs_j;
    const int64_t n_batches = (m + bs_j - 1) / bs_j;
    scaled_t* bsum = malloc(sizeof(scaled_t) * (size_t)n_batches);
    if (!bsum) return -1;
    for (int k = 0; k < D->n_reps; ++k) {
#pragma omp parallel num_threads(D->threads > 0 ? D->threads : omp_get_max_threads())
        {
            double* Y = malloc(sizeof(double) * (size_t)bs_j * (size_t)(r > 1 ? r - 1 : 1));
            double* cv = malloc(sizeof(double) * (size_t)bs_j);
            double* dcv = malloc(sizeof(double) * (size_t)bs_j);
            double* ebv = malloc(sizeof(double) * (size_t)bs_j);
            int* mdv = malloc(sizeof(int) * (size_t)bs_j);
            double* pm = malloc(sizeof(double) * (size_t)bs_j);
            int* pe = malloc(sizeof(int) * (size_t)bs_j);
#pragma omp for schedule(static)
            for (int64_t bi = 0; bi < n_batches; ++bi) {
                const int64_t j1 = bi * bs_j + 1;
                const int64_t j2 = j1 + bs_j - 1 < m ? j1 + bs_j - 1 : m;
                const int len = (int)(j2 - j1 + 1);
                const scaled_t p0 = sc_make(dc_1);
                for (int t = 0; t < len; ++t) {
                    cv[t] = c_1;
                    dcv[t] = dc_1;
                    ebv[t] = eb_1;
                    mdv[t] = mode_1;
                    pm[t] = p0.m;
                    pe[t] = p0.e;
                }
                for (int g = 0; g < r - 1; ++g) {
                    const int i = g + 1;
                    for (int t = 0; t < len; ++t) {
                        double* Yt = Y + (size_t)t * (size_t)(r - 1);
                        Yt[g] = draw_y(D, P, g + 1, k, j1 + t, cv[t], dcv[t], ebv[t], mdv[t]);
                        double za = -INFINITY, zb = INFINITY;
                        for (int c = beg[i]; c < beg[i + 1]; ++c) {
                            const double* uc = U + (size_t)c * n;
                            double sc = 0.0;
                            for (int l = 0; l < i; ++l) sc = fma(Yt[l], uc[l], sc);
                            za = fmax(za, (a[c] - sc) * inv[c]);
                            zb = fmin(zb, (b[c] - sc) * inv[c]);
                        }
                        if (!(za <= zb)) zb = za;
                        phi_state_z(D->phi_form, za, zb, &cv[t], &dcv[t], &ebv[t], &mdv[t]);
                        int e;
                        pm[t] = frexp(pm[t] * dcv[t], &e);
                        pe[t] += e;
                    }
                }
                bsum[bi] = batch_sum(pm, pe, len);
            }
            free(Y); free(cv); free(dcv); free(ebv); free(mdv); free(pm); free(pe);
        }
        scaled_t tot = {0.0, 0};
        for (int64_t bi = 0; bi < n_batches; ++bi) tot = sc_add(tot, bsum[bi]);
        scaled_t rep = {rep_m[k], rep_e[k]};
        rep = rep_merge(rep, tot, n_pts0, m);
        rep_m[k] = rep.m;
        rep_e[k] = rep.e;
    }
    free(bsum);
    return 0;
}

/* Sparse qmc_loop!  src/qmc.jl:618-718.  colptr/rowval/nzval: CSC of
 * triu(U, 1) with 0-based pointers and rows (rows ascending per column);
 * Ud: diagonal of U.  tent = 1 follows src/qmc.jl; tent = 0 with Sobol
 * follows SparseQMC.qmc_loop! (src/qmc_sparse.jl:195-207: no tent, c + y*dc,
 * reductions with plain division, :254-260). */
int orc_qmc_loop_sparse(const orc_desc* D, const orc_points* P, const int64_t* colptr,
                        const int32_t* rowval, const double* nzval, const double* Ud,
                        const double* a, const double* b, int64_t n_pts0, double c_1,
                        double dc_1, double eb_1, int mode_1, int tent, double* rep_m,
                        int* rep_e)
{
    const int n = D->n;
    const int64_t m = D->m;
    const int bs_j = D->bs_j;
    const double c2 = sqrt(2.0);
    const int64_t n_batches = (m + bs_j - 1) / bs_j;
    scaled_t* bsum = malloc(sizeof(scaled_t) * (size_t)n_batches);
    if (!bsum) return -1;
    const int nd = n - 1;
    for (int k = 0; k < D->n_reps; ++k) {
#pragma omp parallel num_threads(D->threads > 0 ? D->threads : omp_get_max_threads())
        {
            double* Y = malloc(sizeof(double) * (size_t)bs_j * (size_t)(nd > 0 ? nd : 1));
            double* cv = malloc(sizeof(double) * (size_t)bs_j);
            double* dcv = malloc(sizeof(double) * (size_t)bs_j);
            double* ebv = malloc(sizeof(double) * (size_t)bs_j);
            int* mdv = malloc(sizeof(int) * (size_t)bs_j);
            double* pm = malloc(sizeof(double) * (size_t)bs_j);
            int* pe = malloc(sizeof(int) * (size_t)bs_j);
#pragma omp for schedule(static)
            for (int64_t bi = 0; bi < n_batches; ++bi) {
                const int64_t j1 = bi * bs_j + 1;
                const int64_t j2 = j1 + bs_j - 1 < m ? j1 + bs_j - 1 : m;
                const int len = (int)(j2 - j1 + 1);
                scaled_t p0 = sc_make(dc_1);
                for (int t = 0; t < bs_j; ++t) {
                    cv[t] = c_1;
                    dcv[t] = dc_1;
                    ebv[t] = eb_1;
                    mdv[t] = mode_1;
                    pm[t] = p0.m;
                    pe[t] = p0.e;
                }
                for (int i = 1; i <= nd; ++i) {
                    for (int t = 0; t < len; ++t) {
                        double yv;
                        if (tent) {
                            Y[(size_t)t + (size_t)(i - 1) * bs_j] = draw_y(D, P, i, k, j1 + t, cv[t], dcv[t], ebv[t], mdv[t]);
                            continue;
                        } else {
                            const int L = D->n_bits;
                            const uint64_t mask = (L >= 64) ? ~0ull : ((1ull << L) - 1);
                            const uint64_t shift_i = P->shifts_u[(size_t)(i - 1) + (size_t)k * nd] & mask;
                            const uint64_t x = (uint64_t)sobol_x(P->V + (size_t)(i - 1) * 32, P->sobol_first + (uint64_t)(j1 + t));
                            const double y_i = (double)((x ^ shift_i) & mask) * ldexp(1.0, -L);
                            const double pp = fma(y_i, dcv[t], cv[t]);
                            yv = fma(2.0, fmin(fmax(pp, EPS_T), 1.0 - EPS_T), -1.0);
                        }
                        Y[(size_t)t + (size_t)(i - 1) * bs_j] = c2 * orc_erfinv(yv);
                    }
                    const int col = i; /* 0-based column i+1-1 */
                    const double a_i = a[i], b_i = b[i];
                    const double inv = 1.0 / (Ud[i] * c2);
                    for (int t = 0; t < len; ++t) {
                        double sv = 0.0;
                        for (int64_t p = colptr[col]; p < colptr[col + 1]; ++p)
                            sv = fma(Y[(size_t)t + (size_t)rowval[p] * bs_j], nzval[p], sv);
                        phi_state(D->phi_form == 2 ? 2 : 0, sv, a_i, b_i, inv, &cv[t], &dcv[t], &ebv[t], &mdv[t]);
                        const double dcval = dcv[t];
                        int e;
                        pm[t] = frexp(pm[t] * dcval, &e);
                        pe[t] += e;
                    }
                }
                bsum[bi] = batch_sum(pm, pe, len);
            }
            free(Y); free(cv); free(dcv); free(ebv); free(mdv); free(pm); free(pe);
        }
        scaled_t tot = {0.0, 0};
        for (int64_t bi = 0; bi < n_batches; ++bi) tot = sc_add(tot, bsum[bi]);
        scaled_t rep = {rep_m[k], rep_e[k]};
        if (tent) {
            rep = rep_merge(rep, tot, n_pts0, m);
        } else {
            /* qmc_reps[k] = qmc_reps[k] * (n0 / n_tot) + sum / n_tot */
            const int64_t n_tot = n_pts0 + m;
            scaled_t mean = tot;
            mean.m = mean.m / (double)n_tot;
            const int E = sc_common_exp(rep, mean);
            rep.m = ldexp(rep.m, rep.e - E) * ((double)n_pts0 / (double)n_tot) + ldexp(mean.m, mean.e - E);
            rep.e = E;
        }
        rep_m[k] = rep.m;
        rep_e[k] = rep.e;
    }
    free(bsum);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* qmc_pnorm!  src/qmc.jl:783-850 (and SparseQMC.qmc_pnorm! :272-317)         */
/* ------------------------------------------------------------------------ */
typedef struct {
    double value;      /* P (0 when it underflows Float64) */
    double log_value;  /* log P */
    double err3;       /* 3 * standard error of the shift means */
    double rel_err3;   /* err3 / value, finite even when both underflow */
    int64_t n_pts;
    int rounds;
} orc_result;

/* result and 3*sd over the shift means (sd: src/qmc.jl:303-312) */
static void reps_summary(const double* rep_m, const int* rep_e, int n_reps, double* mean_m,
                         double* err_m, int* E_out)
{
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
    *mean_m = result;
    *err_m = 3.0 * sqrt(acc / (n_reps - 1));
    *E_out = E;
}

/* kind: 0 dense, 1 sparse (qmc.jl form), 2 SparseQMC module form.
 * gen_state: in/out generator block state.  Dense/sparse: index r of the
 * current block (Richtmyer n_base = r*m; Sobol first index = skip + r*m);
 * next_points! increments it.  SparseQMC: number of blocks generated so far
 * (round 0 always re-uses block 0, copyto!(D.X, D.X0)). */
/* the level structure of a rank-deficient plan (kind 3) */
typedef struct {
    const double* inv;
    const int* beg;
    int r;
} lowrank_t;

static int pnorm_rounds(const orc_desc* D, int kind, const double* U, const double* a, const double* b,
                        const int64_t* colptr, const int32_t* rowval, const double* nzval, const lowrank_t* LR,
                        double c_1, double dc_1, double eb_1, int mode_1, const double* q,
                        const double* shifts_f, const uint32_t* V, const uint64_t* shifts_u,
                        uint64_t sobol_skip, int64_t max_pts, double max_abs_err, int64_t* gen_state,
                        double* rep_m, int* rep_e, orc_result* out);

int orc_qmc_pnorm(const orc_desc* D, int kind, const double* U, const double* a, const double* b,
                  const int64_t* colptr, const int32_t* rowval, const double* nzval,
                  const double* q, const double* shifts_f, const uint32_t* V,
                  const uint64_t* shifts_u, uint64_t sobol_skip, int64_t max_pts,
                  double max_abs_err, int64_t* gen_state, double* rep_m, int* rep_e,
                  orc_result* out)
{
    const double u11 = U[0];
    const double sq2 = sqrt(2.0);
    const double arg_a = a[0] / (u11 * sq2);
    const double arg_b = b[0] / (u11 * sq2);
    double c_1, dc_1, eb_1 = 0.0;
    int mode_1 = 0;
    first_state(D->phi_form, kind, arg_a, arg_b, &c_1, &dc_1, &eb_1, &mode_1);
    return pnorm_rounds(D, kind, U, a, b, colptr, rowval, nzval, NULL, c_1, dc_1, eb_1, mode_1, q, shifts_f, V,
                        shifts_u, sobol_skip, max_pts, max_abs_err, gen_state, rep_m, rep_e, out);
}

/* qmc_pnorm! (src/qmc.jl:783-850) over the level structure of a rank-
 * deficient factor (orc_qmc_loop_lowrank): U, a, b, inv, beg as there,
 * r >= 1 latent variables, r - 1 sampled dimensions. */
int orc_qmc_pnorm_lowrank(const orc_desc* D, const double* U, const double* a, const double* b,
                          const double* inv, const int* beg, int r, const double* q, const double* shifts_f,
                          const uint32_t* V, const uint64_t* shifts_u, uint64_t sobol_skip, int64_t max_pts,
                          double max_abs_err, int64_t* gen_state, double* rep_m, int* rep_e, orc_result* out)
{
    double c_1, dc_1, eb_1;
    int mode_1;
    lowrank_first(D->phi_form, U, D->n, a, b, beg, r, &c_1, &dc_1, &eb_1, &mode_1);
    const lowrank_t LR = {inv, beg, r};
    return pnorm_rounds(D, 3, U, a, b, NULL, NULL, NULL, &LR, c_1, dc_1, eb_1, mode_1, q, shifts_f, V, shifts_u,
                        sobol_skip, max_pts, max_abs_err, gen_state, rep_m, rep_e, out);
}

static int pnorm_rounds(const orc_desc* D, int kind, const double* U, const double* a, const double* b,
                        const int64_t* colptr, const int32_t* rowval, const double* nzval, const lowrank_t* LR,
                        double c_1, double dc_1, double eb_1, int mode_1, const double* q,
                        const double* shifts_f, const uint32_t* V, const uint64_t* shifts_u,
                        uint64_t sobol_skip, int64_t max_pts, double max_abs_err, int64_t* gen_state,
                        double* rep_m, int* rep_e, orc_result* out)
{
    const int n = D->n;
    double* Ud = NULL;
    if (kind == 1 || kind == 2) {
        Ud = malloc(sizeof(double) * (size_t)n);
        for (int i = 0; i < n; ++i) Ud[i] = U[(size_t)i + (size_t)i * n];
    }
    for (int k = 0; k < D->n_reps; ++k) {
        rep_m[k] = 0.0;
        rep_e[k] = 0;
    }
    orc_points P;
    memset(&P, 0, sizeof(P));
    P.q = q;
    P.shifts_f = shifts_f;
    P.V = V;
    P.shifts_u = shifts_u;
    int64_t n_pts = 0;
    int rounds = 0;
    double mean_m, err_m;
    int E;
    for (;;) {
        int64_t block;
        if (kind == 2)
            block = rounds == 0 ? 0 : (*gen_state)++;
        else
            block = *gen_state;
        P.n_base = block * D->m;
        P.sobol_first = sobol_skip + (uint64_t)block * (uint64_t)D->m;
        if (n == 1 || (kind == 3 && LR->r == 1)) {
            /* no sampled dimension: every point has p = dc_1 */
            for (int k = 0; k < D->n_reps; ++k) {
                scaled_t pz = sc_make(dc_1);
                scaled_t tot = pz;
                tot.m *= (double)D->m;
                scaled_t rep = {rep_m[k], rep_e[k]};
                rep = rep_merge(rep, sc_make(sc_to_double(tot)), n_pts, D->m);
                rep_m[k] = rep.m;
                rep_e[k] = rep.e;
            }
        } else if (kind == 0) {
            orc_qmc_loop_dense(D, &P, U, a, b, n_pts, c_1, dc_1, eb_1, mode_1, rep_m, rep_e);
        } else if (kind == 3) {
            orc_qmc_loop_lowrank(D, &P, U, a, b, LR->inv, LR->beg, LR->r, n_pts, c_1, dc_1, eb_1, mode_1, rep_m,
                                 rep_e);
        } else {
            orc_qmc_loop_sparse(D, &P, colptr, rowval, nzval, Ud, a, b, n_pts, c_1, dc_1, eb_1,
                                mode_1, kind == 1, rep_m, rep_e);
        }
        rounds++;
        reps_summary(rep_m, rep_e, D->n_reps, &mean_m, &err_m, &E);
        n_pts += D->m;
        const double err_acc = ldexp(err_m, E);
        if (rounds == 1 && ((n_pts >= max_pts) || (err_acc < max_abs_err))) break;
        if (!((n_pts < max_pts) && (err_acc > max_abs_err))) break;
        if (kind != 2) (*gen_state)++; /* next_points! */
    }
    free(Ud);
    out->value = ldexp(mean_m, E);
    out->log_value = mean_m > 0 ? log(mean_m) + E * M_LN2 : (mean_m == 0 ? -INFINITY : NAN);
    out->err3 = ldexp(err_m, E);
    out->rel_err3 = err_m / mean_m;
    out->n_pts = n_pts;
    out->rounds = rounds;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Adjoint of the tail-form estimate (no reference counterpart)              */
/* ------------------------------------------------------------------------ */
/* For one sample with points w_i, the integrand is p = prod_i dc_i with
 *   alpha_i = (a_i - s_i) / U_ii, beta_i = (b_i - s_i) / U_ii, s_0 = 0,
 *   c_i = Phi(alpha_i), dc_i = Phi(beta_i) - Phi(alpha_i),
 *   y_i = Phi^{-1}(c_i + w_i dc_i)             (i <= d - 2),
 *   s_t = sum_{l < t} U[l, t] y_l.
 * log P = log mean(p), so d log P = sum_s p_s dL_s / sum_s p_s with
 * L = sum_i log dc_i.  The reverse sweep, i = d-1 .. 0:
 *   dcbar_i = 1/dc_i + ybar_i w_i / phi(y_i),  cbar_i = ybar_i / phi(y_i)
 *   albar_i = (cbar_i - dcbar_i) phi(alpha_i), bebar_i = dcbar_i phi(beta_i)
 *   abar_i += albar_i / U_ii, bbar_i += bebar_i / U_ii,
 *   sbar_i = -(albar_i + bebar_i) / U_ii,
 *   Ubar_ii += -(albar_i alpha_i + bebar_i beta_i) / U_ii,
 *   ybar_l += U[l, i] sbar_i, Ubar[l, i] += sbar_i y_l   (l < i).
 * Phi^{-1} is differentiated as a function (dy/du = 1/phi(y)); the tail
 * form only changes how u is evaluated, not the function.  The weights p_s
 * are carried relative to a running maximum of L (rescaled when it grows). */
static inline double npdf(double x) { return isinf(x) ? 0.0 : exp(-0.5 * x * x) * 0.39894228040143267794; }

typedef struct {
    double* ga;
    double* gb;
    double* gU;
    double W;     /* sum of weights exp(L - M) */
    double M;     /* running maximum of L */
} grad_acc;

static void grad_rescale(grad_acc* G, int n, double M_new)
{
    const double f = exp(G->M - M_new);
    for (int i = 0; i < n; ++i) {
        G->ga[i] *= f;
        G->gb[i] *= f;
    }
    for (size_t e = 0; e < (size_t)n * n; ++e) G->gU[e] *= f;
    G->W *= f;
    G->M = M_new;
}

int orc_qmc_grad_dense(const orc_desc* D, const double* U, const double* a, const double* b,
                       const double* q, const double* shifts_f, const uint32_t* V,
                       const uint64_t* shifts_u, uint64_t sobol_skip, double* ga, double* gb,
                       double* gU, double* logP)
{
    const int n = D->n, nd = n - 1;
    const double sq2 = sqrt(2.0);
    orc_points P;
    memset(&P, 0, sizeof(P));
    P.q = q;
    P.shifts_f = shifts_f;
    P.V = V;
    P.shifts_u = shifts_u;
    P.sobol_first = sobol_skip;
    const int64_t total = D->m * (int64_t)D->n_reps;
    int nt = 1;
#ifdef _OPENMP
    nt = D->threads > 0 ? D->threads : omp_get_max_threads();
#endif
    grad_acc* acc = calloc((size_t)nt, sizeof(grad_acc));
    for (int t = 0; t < nt; ++t) {
        acc[t].ga = calloc((size_t)n, sizeof(double));
        acc[t].gb = calloc((size_t)n, sizeof(double));
        acc[t].gU = calloc((size_t)n * n, sizeof(double));
        acc[t].M = -INFINITY;
    }
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        grad_acc* G = &acc[tid];
        double* al = malloc(sizeof(double) * (size_t)n);
        double* be = malloc(sizeof(double) * (size_t)n);
        double* dc = malloc(sizeof(double) * (size_t)n);
        double* w = malloc(sizeof(double) * (size_t)n);
        double* y = malloc(sizeof(double) * (size_t)n);
        double* yb = malloc(sizeof(double) * (size_t)n);
        double* lga = malloc(sizeof(double) * (size_t)n);
        double* lgb = malloc(sizeof(double) * (size_t)n);
        double* lgd = malloc(sizeof(double) * (size_t)n);   /* diagonal of Ubar */
        double* sb = malloc(sizeof(double) * (size_t)n);
#pragma omp for schedule(static)
        for (int64_t smp = 0; smp < total; ++smp) {
            const int k = (int)(smp / D->m);
            const int64_t j = smp % D->m + 1;
            /* forward */
            double L = 0.0;
            double s = 0.0;
            for (int i = 0; i < n; ++i) {
                if (i > 0) {
                    s = 0.0;
                    for (int l = 0; l < i; ++l) s += U[(size_t)l + (size_t)i * n] * y[l];
                }
                const double uii = U[(size_t)i + (size_t)i * n];
                al[i] = (a[i] - s) / uii;
                be[i] = (b[i] - s) / uii;
                double ea, eb, dci;
                int mode;
                tail_masses(al[i] / sq2, be[i] / sq2, &ea, &eb, &dci, &mode);
                dc[i] = dci;
                L += log(dci);
                if (i < nd) {
                    w[i] = fabs(fma(point_u(D, &P, i + 1, k, j), 2.0, -1.0));
                    y[i] = tail_sample(w[i], ea, eb, dci, mode);
                }
            }
            if (!(L > -INFINITY)) continue;  /* zero weight */
            /* reverse */
            for (int l = 0; l < n; ++l) {
                yb[l] = 0.0;
                sb[l] = 0.0;
            }
            for (int i = n - 1; i >= 0; --i) {
                double dcb = 1.0 / dc[i], cb = 0.0;
                if (i < nd) {
                    const double iphy = 1.0 / npdf(y[i]);
                    dcb += yb[i] * w[i] * iphy;
                    cb = yb[i] * iphy;
                }
                const double pa = npdf(al[i]), pb = npdf(be[i]);
                const double alb = (cb - dcb) * pa, beb = dcb * pb;
                const double uii = U[(size_t)i + (size_t)i * n];
                lga[i] = alb / uii;
                lgb[i] = beb / uii;
                const double ta = isinf(al[i]) ? 0.0 : alb * al[i];
                const double tb = isinf(be[i]) ? 0.0 : beb * be[i];
                lgd[i] = -(ta + tb) / uii;
                sb[i] = -(alb + beb) / uii;
                if (i > 0)
                    for (int l = 0; l < i; ++l) yb[l] += U[(size_t)l + (size_t)i * n] * sb[i];
            }
            /* accumulate with weight exp(L - M) */
            if (L > G->M) grad_rescale(G, n, L);
            const double wt = exp(L - G->M);
            G->W += wt;
            for (int i = 0; i < n; ++i) {
                G->ga[i] += wt * lga[i];
                G->gb[i] += wt * lgb[i];
                G->gU[(size_t)i + (size_t)i * n] += wt * lgd[i];
                for (int l = 0; l < i; ++l) G->gU[(size_t)l + (size_t)i * n] += wt * sb[i] * y[l];
            }
        }
        free(al); free(be); free(dc); free(w); free(y); free(yb); free(lga); free(lgb); free(lgd); free(sb);
    }
    /* merge the threads */
    double M = -INFINITY;
    for (int t = 0; t < nt; ++t) M = fmax(M, acc[t].M);
    double W = 0.0;
    for (int i = 0; i < n; ++i) ga[i] = gb[i] = 0.0;
    for (size_t e = 0; e < (size_t)n * n; ++e) gU[e] = 0.0;
    for (int t = 0; t < nt; ++t) {
        if (acc[t].M == -INFINITY) continue;
        grad_rescale(&acc[t], n, M);
        W += acc[t].W;
        for (int i = 0; i < n; ++i) {
            ga[i] += acc[t].ga[i];
            gb[i] += acc[t].gb[i];
        }
        for (size_t e = 0; e < (size_t)n * n; ++e) gU[e] += acc[t].gU[e];
    }
    for (int i = 0; i < n; ++i) {
        ga[i] /= W;
        gb[i] /= W;
    }
    for (size_t e = 0; e < (size_t)n * n; ++e) gU[e] /= W;
    /* log P = M + log(W / total) */
    *logP = M + log(W / (double)total);
    for (int t = 0; t < nt; ++t) {
        free(acc[t].ga);
        free(acc[t].gb);
        free(acc[t].gU);
    }
    free(acc);
    return 0;
}

/* The batch loop uses orc_erfinv; expose erf-form helpers for the tests. */
double orc_phi_c(double x) { return 0.5 * (1.0 + erf(x)); }

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
