/*
 * oracle.c -- CPU restatement of the reference's CPU path (netlib LAPACK looped
 * over the batch under OpenMP).  TEST INFRASTRUCTURE ONLY: loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 *
 * Each routine cites the netlib algorithm it restates and the reference call
 * site that uses it.  See oracle.h for the rounding contract.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define A_(i, j) A[(size_t)(i) + (size_t)(j) * (size_t)lda]
#define B_(i, j) B[(size_t)(i) + (size_t)(j) * (size_t)ldb]
#define C_(i, j) C[(size_t)(i) + (size_t)(j) * (size_t)ldc]

static int imin(int a, int b) { return a < b ? a : b; }

double orc_eps(void) { return ldexp(1.0, -53); }

double orc_wtime(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* netlib BLAS idamax: first index of the maximum |x|, comparisons strict, so a
 * NaN is selected only when it is the first element.  0-based result.
 * Used by dgetf2/dgetrf2 (pivot search); MAGMA's GPU equivalent is
 * magmablas/zgetf2_kernels.cu:37-72. */
int orc_idamax(int n, const double *x)
{
    if (n < 1) return -1;
    int best = 0;
    double dmax = fabs(x[0]);
    for (int i = 1; i < n; ++i) {
        double v = fabs(x[i]);
        if (v > dmax) { best = i; dmax = v; }
    }
    return best;
}

/* netlib dgetf2 (right-looking unblocked LU with partial pivoting).
 * Reference call site: lapackf77_dgetrf in testing/testing_zgetrf_batched.cpp:200
 * and testing/testing_zgesv_batched.cpp:171.  netlib dgetrf/dgetrf2 with
 * reference BLAS (dtrsm, dgemm, dger loop orders) apply to every element the
 * same updates in the same order as dgetf2, so this one loop nest stands for
 * all three.  Multiplier scaling uses the reciprocal when |pivot| >= sfmin,
 * division otherwise (dgetf2.f / dgetrf2.f).  A zero pivot sets info (first
 * one only) and the factorization continues. */
void orc_dgetrf(int m, int n, double *A, int lda, int *ipiv, int *info)
{
    *info = 0;
    const int mn = imin(m, n);
    const double sfmin = DBL_MIN;          /* dlamch('S') */
    for (int j = 0; j < mn; ++j) {
        const int jp = j + orc_idamax(m - j, &A_(j, j));
        ipiv[j] = jp + 1;
        if (A_(jp, j) != 0.0) {
            if (jp != j) {
                for (int c = 0; c < n; ++c) {
                    double t = A_(j, c); A_(j, c) = A_(jp, c); A_(jp, c) = t;
                }
            }
            if (j < m - 1) {
                const double p = A_(j, j);
                if (fabs(p) >= sfmin) {
                    const double r = 1.0 / p;
                    for (int i = j + 1; i < m; ++i) A_(i, j) = r * A_(i, j);
                } else {
                    for (int i = j + 1; i < m; ++i) A_(i, j) = A_(i, j) / p;
                }
            }
        } else if (*info == 0) {
            *info = j + 1;
        }
        if (j < mn - 1) {
            /* dger: A22 := A22 - x*y', column by column, skipping y_c == 0 */
            for (int c = j + 1; c < n; ++c) {
                const double y = A_(j, c);
                if (y != 0.0) {
                    const double t = -y;
                    for (int i = j + 1; i < m; ++i) A_(i, c) = fma(A_(i, j), t, A_(i, c));
                }
            }
        }
    }
}

/* LU without pivoting: orc_dgetrf with the pivot search removed (pivot =
 * the diagonal entry), the same multiplier rule, the same dger zero skip and
 * the same info rule (first zero pivot; the factorization continues with the
 * column unscaled).  Semantics of magma_dgetrf_nopiv_batched
 * (src/zgetrf_nopiv_batched.cpp:76-306); netlib LAPACK has no such routine. */
void orc_dgetrf_nopiv(int m, int n, double *A, int lda, int *info)
{
    *info = 0;
    const int mn = imin(m, n);
    const double sfmin = DBL_MIN;
    for (int j = 0; j < mn; ++j) {
        const double p = A_(j, j);
        if (p != 0.0) {
            if (fabs(p) >= sfmin) {
                const double r = 1.0 / p;
                for (int i = j + 1; i < m; ++i) A_(i, j) = r * A_(i, j);
            } else {
                for (int i = j + 1; i < m; ++i) A_(i, j) = A_(i, j) / p;
            }
        } else if (*info == 0) {
            *info = j + 1;
        }
        if (j < mn - 1) {
            for (int c = j + 1; c < n; ++c) {
                const double y = A_(j, c);
                if (y != 0.0) {
                    const double t = -y;
                    for (int i = j + 1; i < m; ++i) A_(i, c) = fma(A_(i, j), t, A_(i, c));
                }
            }
        }
    }
}

/* netlib dlaswp: row interchanges k1..k2 (1-based, inclusive) applied in order
 * (incx > 0) or in reverse order (incx < 0).  Used by dgetrs and by the tester's
 * LU check (testing/testing_zgetrf_batched.cpp:47). */
void orc_dlaswp(int ncol, double *A, int lda, int k1, int k2, const int *ipiv, int incx)
{
    if (incx > 0) {
        for (int i = k1; i <= k2; ++i) {
            int ip = ipiv[i - 1];
            if (ip != i)
                for (int c = 0; c < ncol; ++c) {
                    double t = A_(i - 1, c); A_(i - 1, c) = A_(ip - 1, c); A_(ip - 1, c) = t;
                }
        }
    } else {
        for (int i = k2; i >= k1; --i) {
            int ip = ipiv[i - 1];
            if (ip != i)
                for (int c = 0; c < ncol; ++c) {
                    double t = A_(i - 1, c); A_(i - 1, c) = A_(ip - 1, c); A_(ip - 1, c) = t;
                }
        }
    }
}

/* reference-BLAS dtrsm, Left side, the four triangle/transpose cases dgetrs and
 * dpotrs use.  Loop orders and the zero-skip / division placement follow
 * dtrsm.f exactly. */
static void trsm_left(char uplo, char trans, char diag, int n, int nrhs,
                      const double *A, int lda, double *B, int ldb)
{
    const int nounit = (diag == 'N');
    for (int j = 0; j < nrhs; ++j) {
        if (trans == 'N') {
            if (uplo == 'U') {                      /* B := inv(U) B, k descending */
                for (int k = n - 1; k >= 0; --k) {
                    if (B_(k, j) != 0.0) {
                        if (nounit) B_(k, j) = B_(k, j) / A_(k, k);
                        const double bk = -B_(k, j);
                        for (int i = 0; i < k; ++i) B_(i, j) = fma(bk, A_(i, k), B_(i, j));
                    }
                }
            } else {                                /* B := inv(L) B, k ascending */
                for (int k = 0; k < n; ++k) {
                    if (B_(k, j) != 0.0) {
                        if (nounit) B_(k, j) = B_(k, j) / A_(k, k);
                        const double bk = -B_(k, j);
                        for (int i = k + 1; i < n; ++i) B_(i, j) = fma(bk, A_(i, k), B_(i, j));
                    }
                }
            }
        } else {
            if (uplo == 'U') {                      /* B := inv(U') B, dot form, i ascending */
                for (int i = 0; i < n; ++i) {
                    double t = B_(i, j);
                    for (int k = 0; k < i; ++k) t = fma(-A_(k, i), B_(k, j), t);
                    if (nounit) t = t / A_(i, i);
                    B_(i, j) = t;
                }
            } else {                                /* B := inv(L') B, dot form, i descending */
                for (int i = n - 1; i >= 0; --i) {
                    double t = B_(i, j);
                    for (int k = i + 1; k < n; ++k) t = fma(-A_(k, i), B_(k, j), t);
                    if (nounit) t = t / A_(i, i);
                    B_(i, j) = t;
                }
            }
        }
    }
}

/* netlib dgetrs.  Reference call site: testing/testing_zgesv_batched.cpp:171
 * (via dgesv).  NoTrans: laswp forward, L (unit) solve, U solve.
 * Trans/ConjTrans: U' solve, L' (unit) solve, laswp backward. */
void orc_dgetrs(char trans, int n, int nrhs, const double *A, int lda,
                const int *ipiv, double *B, int ldb)
{
    if (n == 0 || nrhs == 0) return;
    if (trans == 'N') {
        orc_dlaswp(nrhs, B, ldb, 1, n, ipiv, 1);
        trsm_left('L', 'N', 'U', n, nrhs, A, lda, B, ldb);
        trsm_left('U', 'N', 'N', n, nrhs, A, lda, B, ldb);
    } else {
        trsm_left('U', 'T', 'N', n, nrhs, A, lda, B, ldb);
        trsm_left('L', 'T', 'U', n, nrhs, A, lda, B, ldb);
        orc_dlaswp(nrhs, B, ldb, 1, n, ipiv, -1);
    }
}

/* netlib dpotrf/dpotrf2, lower: element (i,j) of L is
 *   (a_ij - l_i0 l_j0 - l_i1 l_j1 - ... ) * (1/l_jj)   (updates in increasing k)
 * exactly as the recursive dpotrf2 + reference dtrsm('R','L','T','N') (reciprocal
 * scaling) + dsyrk('L','N') produce it; the diagonal is sqrt of the updated a_jj.
 * info = first j with updated a_jj <= 0 or NaN; the routine then returns.
 * Reference call site: testing/testing_zpotrf_batched.cpp:137.
 * Upper: MAGMA's batched potrf supports Lower only (magmablas/zpotf2_kernels.cu:153-156
 * rejects Upper); the oracle restates Upper as the mirror image for the
 * solve tests (dpotrs with uplo='U' uses a factor built here). */
void orc_dpotrf(char uplo, int n, double *A, int lda, int *info)
{
    *info = 0;
    const int lower = (uplo == 'L');
    for (int j = 0; j < n; ++j) {
        double ajj = A_(j, j);
        if (!(ajj > 0.0)) { *info = j + 1; return; }
        ajj = sqrt(ajj);
        A_(j, j) = ajj;
        const double r = 1.0 / ajj;
        if (lower) {
            for (int i = j + 1; i < n; ++i) A_(i, j) = r * A_(i, j);
            for (int c = j + 1; c < n; ++c) {
                const double t = A_(c, j);
                if (t != 0.0) {
                    const double nt = -t;
                    for (int i = c; i < n; ++i) A_(i, c) = fma(nt, A_(i, j), A_(i, c));
                }
            }
        } else {
            for (int i = j + 1; i < n; ++i) A_(j, i) = r * A_(j, i);
            for (int c = j + 1; c < n; ++c) {
                const double t = A_(j, c);
                if (t != 0.0) {
                    const double nt = -t;
                    for (int i = c; i < n; ++i) A_(c, i) = fma(nt, A_(j, i), A_(c, i));
                }
            }
        }
    }
}

/* netlib dpotrs.  Lower: L solve (dtrsm L,L,N,N) then L' solve (L,L,T,N).
 * Upper: U' solve then U solve.  Reference call site:
 * testing/testing_zposv_batched.cpp (lapackf77_zposv). */
void orc_dpotrs(char uplo, int n, int nrhs, const double *A, int lda, double *B, int ldb)
{
    if (n == 0 || nrhs == 0) return;
    if (uplo == 'L') {
        trsm_left('L', 'N', 'N', n, nrhs, A, lda, B, ldb);
        trsm_left('L', 'T', 'N', n, nrhs, A, lda, B, ldb);
    } else {
        trsm_left('U', 'T', 'N', n, nrhs, A, lda, B, ldb);
        trsm_left('U', 'N', 'N', n, nrhs, A, lda, B, ldb);
    }
}

/* reference-BLAS dgemm loop orders (dgemm.f): C := alpha*op(A)*op(B) + beta*C.
 * Reference call site: testing/testing_zgemm_vbatched.cpp:273-305. */
void orc_dgemm(char transa, char transb, int m, int n, int k, double alpha,
               const double *A, int lda, const double *B, int ldb,
               double beta, double *C, int ldc)
{
    if (m == 0 || n == 0) return;
    const int nota = (transa == 'N'), notb = (transb == 'N');
    for (int j = 0; j < n; ++j) {
        if (nota) {
            if (beta == 0.0)      for (int i = 0; i < m; ++i) C_(i, j) = 0.0;
            else if (beta != 1.0) for (int i = 0; i < m; ++i) C_(i, j) = beta * C_(i, j);
            for (int l = 0; l < k; ++l) {
                const double t = alpha * (notb ? B_(l, j) : B_(j, l));
                for (int i = 0; i < m; ++i) C_(i, j) = fma(t, A_(i, l), C_(i, j));
            }
        } else {
            for (int i = 0; i < m; ++i) {
                double t = 0.0;
                for (int l = 0; l < k; ++l) t = fma(A_(l, i), notb ? B_(l, j) : B_(j, l), t);
                C_(i, j) = (beta == 0.0) ? alpha * t : alpha * t + beta * C_(i, j);
            }
        }
    }
}

/* ---- random butterfly transformation (RBT), depth 2 ------------------------
 * A <- U' A V and b <- U' b with U = U1 U2, V = V1 V2 products of butterfly
 * matrices, as in magmablas/zgerbt_func_batched.cu: level 1 acts on the two
 * halves (entries n.. of u and v), level 2 on the whole order (entries 0..).
 * One butterfly of order N with weights w pairs index x with x + N/2
 * (x < N/2): (r_x, r_{x+N/2}) <- (w_x (r_x + r_{x+N/2}), w_{x+N/2} (r_x - r_{x+N/2})),
 * an odd leftover index unchanged.  Each level is one pass over the rows
 * (every column) and one over the columns (every row).  The reference's
 * kernels update only the paired square, which is a consistent transform only
 * when both orders are multiples of 4; this form is exact for every n. */
static void rbt_rows(int N, int r0, double *A, int lda, int ncols, const double *u)
{
    const int h = N / 2;
    for (int c = 0; c < ncols; ++c) {
        for (int x = 0; x < h; ++x) {
            const double a = A_(r0 + x, c), b = A_(r0 + x + h, c);
            A_(r0 + x, c) = u[x] * (a + b);
            A_(r0 + x + h, c) = u[x + h] * (a - b);
        }
    }
}

static void rbt_cols(int N, int c0, double *A, int lda, int nrows, const double *v)
{
    const int h = N / 2;
    for (int y = 0; y < h; ++y) {
        for (int r = 0; r < nrows; ++r) {
            const double a = A_(r, c0 + y), b = A_(r, c0 + y + h);
            A_(r, c0 + y) = v[y] * (a + b);
            A_(r, c0 + y + h) = v[y + h] * (a - b);
        }
    }
}

/* b <- U' b, one butterfly (the row pass on a single column) */
static void rbt_mtv(int n, const double *u, double *b)
{
    const int h = n / 2;
    for (int x = 0; x < h; ++x) {
        const double a1 = b[x] + b[x + h], a2 = b[x] - b[x + h];
        b[x] = u[x] * a1;
        b[x + h] = u[x + h] * a2;
    }
}

/* x <- V y, one butterfly (magmablas/zgerbt_kernels.cu apply_vector form) */
static void rbt_mv(int n, const double *v, double *b)
{
    const int h = n / 2;
    for (int x = 0; x < h; ++x) {
        const double a1 = v[x] * b[x], a2 = v[x + h] * b[x + h];
        b[x] = a1 + a2;
        b[x + h] = a1 - a2;
    }
}

/* The butterfly entries: 2n values each of u and v, exp((r - 1/2) / 10) with
 * r in [0,1) (the distribution of src/zgerbt_gpu.cpp:16-28), r from
 * splitmix64 of seed 2^32 + 2 i (u) and 2^32 + 2 i + 1 (v): a fixed
 * function of n, so every call and every matrix uses the same transform. */
static uint64_t rbt_mix(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void orc_rbt_butterflies(int n, double *u, double *v)
{
    for (int i = 0; i < 2 * n; ++i) {
        const double ru = (double)(rbt_mix((1ull << 32) + 2ull * i) >> 11) * 0x1.0p-53;
        const double rv = (double)(rbt_mix((1ull << 32) + 2ull * i + 1) >> 11) * 0x1.0p-53;
        u[i] = exp((ru - 0.5) / 10.0);
        v[i] = exp((rv - 0.5) / 10.0);
    }
}

/* dgesv_rbt: A <- U' A V, B <- U' B (every column), LU without pivoting,
 * the two triangular solves, X = V Y in B.  A and B as magma_dgesv_rbt_batched
 * (src/zgesv_rbt_batched.cpp) leaves them. */
void orc_dgesv_rbt(int n, int nrhs, double *A, int lda, double *B, int ldb, int *info)
{
    *info = 0;
    if (n == 0) return;
    double *u = (double *)malloc(sizeof(double) * 4 * (size_t)n);
    double *v = u + 2 * n;
    orc_rbt_butterflies(n, u, v);
    const int h = n / 2;
    rbt_rows(h, 0, A, lda, n, u + n);
    rbt_rows(h, h, A, lda, n, u + n + h);
    rbt_cols(h, 0, A, lda, n, v + n);
    rbt_cols(h, h, A, lda, n, v + n + h);
    rbt_rows(n, 0, A, lda, n, u);
    rbt_cols(n, 0, A, lda, n, v);
    for (int j = 0; j < nrhs; ++j) {
        double *b = &B_(0, j);
        rbt_mtv(h, u + n, b);
        rbt_mtv(h, u + n + h, b + h);
        rbt_mtv(n, u, b);
    }
    orc_dgetrf_nopiv(n, n, A, lda, info);
    trsm_left('L', 'N', 'U', n, nrhs, A, lda, B, ldb);
    trsm_left('U', 'N', 'N', n, nrhs, A, lda, B, ldb);
    for (int j = 0; j < nrhs; ++j) {
        double *b = &B_(0, j);
        rbt_mv(n, v, b);
        rbt_mv(h, v + n, b);
        rbt_mv(h, v + n + h, b + h);
    }
    free(u);
}

/* reference-BLAS dtrsm (dtrsm.f), all eight side/uplo/transpose cases:
 *   Left:  B := alpha * inv(op(A)) * B,   Right: B := alpha * B * inv(op(A)).
 * Loop orders, the zero skips, the alpha placement, and division (Left) versus
 * multiplication by the reciprocal (Right) follow dtrsm.f; every update is one
 * fma.  Reference call site: testing/testing_ztrsm_batched.cpp:234-240. */
void orc_dtrsm(char side, char uplo, char transa, char diag, int m, int n, double alpha,
               const double *A, int lda, double *B, int ldb)
{
    if (m == 0 || n == 0) return;
    const int nounit = (diag == 'N');
    if (alpha == 0.0) {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) B_(i, j) = 0.0;
        return;
    }
    if (side == 'L') {
        if (transa == 'N') {
            for (int j = 0; j < n; ++j) {
                if (alpha != 1.0)
                    for (int i = 0; i < m; ++i) B_(i, j) = alpha * B_(i, j);
                if (uplo == 'U') {
                    for (int k = m - 1; k >= 0; --k) {
                        if (B_(k, j) != 0.0) {
                            if (nounit) B_(k, j) = B_(k, j) / A_(k, k);
                            const double bk = -B_(k, j);
                            for (int i = 0; i < k; ++i) B_(i, j) = fma(bk, A_(i, k), B_(i, j));
                        }
                    }
                } else {
                    for (int k = 0; k < m; ++k) {
                        if (B_(k, j) != 0.0) {
                            if (nounit) B_(k, j) = B_(k, j) / A_(k, k);
                            const double bk = -B_(k, j);
                            for (int i = k + 1; i < m; ++i) B_(i, j) = fma(bk, A_(i, k), B_(i, j));
                        }
                    }
                }
            }
        } else {
            for (int j = 0; j < n; ++j) {
                if (uplo == 'U') {
                    for (int i = 0; i < m; ++i) {
                        double t = alpha * B_(i, j);
                        for (int k = 0; k < i; ++k) t = fma(-A_(k, i), B_(k, j), t);
                        if (nounit) t = t / A_(i, i);
                        B_(i, j) = t;
                    }
                } else {
                    for (int i = m - 1; i >= 0; --i) {
                        double t = alpha * B_(i, j);
                        for (int k = i + 1; k < m; ++k) t = fma(-A_(k, i), B_(k, j), t);
                        if (nounit) t = t / A_(i, i);
                        B_(i, j) = t;
                    }
                }
            }
        }
    } else if (transa == 'N') {
        if (uplo == 'U') {
            for (int j = 0; j < n; ++j) {
                if (alpha != 1.0)
                    for (int i = 0; i < m; ++i) B_(i, j) = alpha * B_(i, j);
                for (int k = 0; k < j; ++k) {
                    if (A_(k, j) != 0.0) {
                        const double ak = -A_(k, j);
                        for (int i = 0; i < m; ++i) B_(i, j) = fma(ak, B_(i, k), B_(i, j));
                    }
                }
                if (nounit) {
                    const double r = 1.0 / A_(j, j);
                    for (int i = 0; i < m; ++i) B_(i, j) = r * B_(i, j);
                }
            }
        } else {
            for (int j = n - 1; j >= 0; --j) {
                if (alpha != 1.0)
                    for (int i = 0; i < m; ++i) B_(i, j) = alpha * B_(i, j);
                for (int k = j + 1; k < n; ++k) {
                    if (A_(k, j) != 0.0) {
                        const double ak = -A_(k, j);
                        for (int i = 0; i < m; ++i) B_(i, j) = fma(ak, B_(i, k), B_(i, j));
                    }
                }
                if (nounit) {
                    const double r = 1.0 / A_(j, j);
                    for (int i = 0; i < m; ++i) B_(i, j) = r * B_(i, j);
                }
            }
        }
    } else {
        if (uplo == 'U') {
            for (int k = n - 1; k >= 0; --k) {
                if (nounit) {
                    const double r = 1.0 / A_(k, k);
                    for (int i = 0; i < m; ++i) B_(i, k) = r * B_(i, k);
                }
                for (int j = 0; j < k; ++j) {
                    if (A_(j, k) != 0.0) {
                        const double ajk = -A_(j, k);
                        for (int i = 0; i < m; ++i) B_(i, j) = fma(ajk, B_(i, k), B_(i, j));
                    }
                }
                if (alpha != 1.0)
                    for (int i = 0; i < m; ++i) B_(i, k) = alpha * B_(i, k);
            }
        } else {
            for (int k = 0; k < n; ++k) {
                if (nounit) {
                    const double r = 1.0 / A_(k, k);
                    for (int i = 0; i < m; ++i) B_(i, k) = r * B_(i, k);
                }
                for (int j = k + 1; j < n; ++j) {
                    if (A_(j, k) != 0.0) {
                        const double ajk = -A_(j, k);
                        for (int i = 0; i < m; ++i) B_(i, j) = fma(ajk, B_(i, k), B_(i, j));
                    }
                }
                if (alpha != 1.0)
                    for (int i = 0; i < m; ++i) B_(i, k) = alpha * B_(i, k);
            }
        }
    }
}

/* reference-BLAS dtrsv (dtrsv.f): x := inv(op(A)) x, increment incx >= 1.
 * NoTrans is the column (axpy) form with the zero skip; Trans is the dot form,
 * with i ascending for Upper and descending for Lower as in dtrsv.f (this
 * last order differs from dtrsm's Left/Trans/Lower).  Every update is one fma.
 * Reference call site: testing/testing_ztrsv_batched.cpp:192-196. */
void orc_dtrsv(char uplo, char trans, char diag, int n, const double *A, int lda, double *x, int incx)
{
    const int nounit = (diag == 'N');
#define X_(i) x[(size_t)(i) * (size_t)incx]
    if (trans == 'N') {
        if (uplo == 'U') {
            for (int j = n - 1; j >= 0; --j) {
                if (X_(j) != 0.0) {
                    if (nounit) X_(j) = X_(j) / A_(j, j);
                    const double t = -X_(j);
                    for (int i = j - 1; i >= 0; --i) X_(i) = fma(t, A_(i, j), X_(i));
                }
            }
        } else {
            for (int j = 0; j < n; ++j) {
                if (X_(j) != 0.0) {
                    if (nounit) X_(j) = X_(j) / A_(j, j);
                    const double t = -X_(j);
                    for (int i = j + 1; i < n; ++i) X_(i) = fma(t, A_(i, j), X_(i));
                }
            }
        }
    } else {
        if (uplo == 'U') {
            for (int j = 0; j < n; ++j) {
                double t = X_(j);
                for (int i = 0; i < j; ++i) t = fma(-A_(i, j), X_(i), t);
                if (nounit) t = t / A_(j, j);
                X_(j) = t;
            }
        } else {
            for (int j = n - 1; j >= 0; --j) {
                double t = X_(j);
                for (int i = n - 1; i > j; --i) t = fma(-A_(i, j), X_(i), t);
                if (nounit) t = t / A_(j, j);
                X_(j) = t;
            }
        }
    }
#undef X_
}

/* reference-BLAS dgemv (dgemv.f, LAPACK 3.x form without the x(j) = 0 skip):
 * y := alpha op(A) x + beta y, increments >= 1.  beta is applied first
 * (beta = 0 sets y to zero, so NaNs in y do not propagate); NoTrans adds
 * (alpha x_j) a_ij column by column, Trans forms each dot product in i order
 * and adds alpha times it.  Every update is one fma.
 * Reference call site: testing/testing_zgemv_batched.cpp:155-159. */
void orc_dgemv(char trans, int m, int n, double alpha, const double *A, int lda,
               const double *x, int incx, double beta, double *y, int incy)
{
    const int leny = (trans == 'N') ? m : n, lenx = (trans == 'N') ? n : m;
    if (m == 0 || n == 0 || (alpha == 0.0 && beta == 1.0)) return;
#define X_(i) x[(size_t)(i) * (size_t)incx]
#define Y_(i) y[(size_t)(i) * (size_t)incy]
    if (beta != 1.0) {
        for (int i = 0; i < leny; ++i) Y_(i) = (beta == 0.0) ? 0.0 : beta * Y_(i);
    }
    if (alpha == 0.0) return;
    if (trans == 'N') {
        for (int j = 0; j < lenx; ++j) {
            const double t = alpha * X_(j);
            for (int i = 0; i < m; ++i) Y_(i) = fma(t, A_(i, j), Y_(i));
        }
    } else {
        for (int j = 0; j < n; ++j) {
            double t = 0.0;
            for (int i = 0; i < m; ++i) t = fma(A_(i, j), X_(i), t);
            Y_(j) = fma(alpha, t, Y_(j));
        }
    }
#undef X_
#undef Y_
}

/* ---- Householder QR (netlib dgeqr2 with dlarfg / dlarf) --------------------
 * Reference call site: lapackf77_zgeqrf in testing/testing_zgeqrf_batched.cpp:224;
 * the tester checks ||R - Q'A||_1 / (n ||A||_1) and ||I - Q'Q||_1 / n
 * (:34-82).  dnrm2 is reference BLAS's scaled sum of squares, dlapy2 the
 * overflow-safe hypotenuse; dlarfg's rescaling of a tiny beta is kept. */
static double nrm2(int n, const double *x)
{
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n; ++i) {
        if (x[i] != 0.0) {
            const double a = fabs(x[i]);
            if (scale < a) {
                ssq = 1.0 + ssq * (scale / a) * (scale / a);
                scale = a;
            } else {
                ssq += (a / scale) * (a / scale);
            }
        }
    }
    return scale * sqrt(ssq);
}

static double lapy2(double x, double y)
{
    const double xa = fabs(x), ya = fabs(y);
    const double w = xa > ya ? xa : ya, z = xa < ya ? xa : ya;
    if (z == 0.0) return w;
    return w * sqrt(1.0 + (z / w) * (z / w));
}

/* dlarfg: H (alpha; x) = (beta; 0), H = I - tau v v', v = (1; x). */
static void larfg(int n, double *alpha, double *x, double *tau)
{
    *tau = 0.0;
    if (n <= 1) return;
    double xnorm = nrm2(n - 1, x);
    if (xnorm == 0.0) return;
    double beta = -copysign(lapy2(*alpha, xnorm), *alpha);
    const double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
    int knt = 0;
    if (fabs(beta) < safmin) {
        const double rsafmn = 1.0 / safmin;
        do {
            ++knt;
            for (int i = 0; i < n - 1; ++i) x[i] *= rsafmn;
            beta *= rsafmn;
            *alpha *= rsafmn;
        } while (fabs(beta) < safmin && knt < 20);
        xnorm = nrm2(n - 1, x);
        beta = -copysign(lapy2(*alpha, xnorm), *alpha);
    }
    *tau = (beta - *alpha) / beta;
    const double sc = 1.0 / (*alpha - beta);
    for (int i = 0; i < n - 1; ++i) x[i] *= sc;
    for (int j = 0; j < knt; ++j) beta *= safmin;
    *alpha = beta;
}

void orc_dgeqrf(int m, int n, double *A, int lda, double *tau)
{
    const int k = imin(m, n);
    double *w = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < k; ++i) {
        larfg(m - i, &A_(i, i), &A_(imin(i + 1, m - 1), i), &tau[i]);
        if (i < n - 1 && tau[i] != 0.0) {
            /* dlarf('Left'): w = C' v, C -= tau v w' with v = (1; A(i+1:m, i)) */
            const double aii = A_(i, i);
            A_(i, i) = 1.0;
            for (int c = i + 1; c < n; ++c) {
                double t = 0.0;
                for (int r = i; r < m; ++r) t = fma(A_(r, c), A_(r, i), t);
                w[c] = t;
            }
            for (int c = i + 1; c < n; ++c) {
                const double t = -tau[i] * w[c];
                for (int r = i; r < m; ++r) A_(r, c) = fma(A_(r, i), t, A_(r, c));
            }
            A_(i, i) = aii;
        }
    }
    free(w);
}

/* LAPACK dlange: 'M' max abs, 'O'/'1' one norm, 'I' infinity norm, 'F' Frobenius. */
double orc_dlange(char norm, int m, int n, const double *A, int lda)
{
    long double v = 0.0L;
    if (m == 0 || n == 0) return 0.0;
    if (norm == 'M') {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) { long double a = fabsl(A_(i, j)); if (a > v || isnan(a)) v = a; }
    } else if (norm == 'O' || norm == '1') {
        for (int j = 0; j < n; ++j) {
            long double s = 0.0L;
            for (int i = 0; i < m; ++i) s += fabsl(A_(i, j));
            if (s > v || isnan(s)) v = s;
        }
    } else if (norm == 'I') {
        for (int i = 0; i < m; ++i) {
            long double s = 0.0L;
            for (int j = 0; j < n; ++j) s += fabsl(A_(i, j));
            if (s > v || isnan(s)) v = s;
        }
    } else {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) v += (long double)A_(i, j) * A_(i, j);
        v = sqrtl(v);
    }
    return (double)v;
}

/* Tester LU check (testing/testing_zgetrf_batched.cpp:30-70):
 *   ||P*A - L*U||_F / (||A||_F * n), returned divided by eps (pass: < 30).
 * Products are accumulated in long double so the check does not add its own
 * FP64 rounding to the quantity it measures. */
double orc_lu_error(int m, int n, const double *A, int lda,
                    const double *LU, int ldlu, const int *ipiv)
{
    const int mn = imin(m, n);
    if (m == 0 || n == 0) return 0.0;
    double *PA = (double *)malloc(sizeof(double) * (size_t)m * n);
    for (int j = 0; j < n; ++j) for (int i = 0; i < m; ++i) PA[i + (size_t)j * m] = A_(i, j);
    orc_dlaswp(n, PA, m, 1, mn, ipiv, 1);
    long double num = 0.0L, den = 0.0L;
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < m; ++i) {
            long double s = 0.0L;
            const int kmax = imin(imin(i, j), mn - 1);
            for (int k = 0; k <= kmax; ++k) {
                long double l = (k == i) ? 1.0L : LU[i + (size_t)k * ldlu];
                long double u = LU[k + (size_t)j * ldlu];
                if (k <= i && k <= j) s += l * u;
            }
            long double d = s - PA[i + (size_t)j * m];
            num += d * d;
            den += (long double)PA[i + (size_t)j * m] * PA[i + (size_t)j * m];
        }
    }
    free(PA);
    if (den == 0.0L) return (num == 0.0L) ? 0.0 : INFINITY;
    return (double)(sqrtl(num) / (sqrtl(den) * n) / orc_eps());
}

/* Tester solve check (testing/testing_zgesv_batched.cpp:133-153):
 *   ||B - A*X||_inf / (n * ||A||_inf * ||X||_inf), returned divided by eps. */
double orc_solve_error(int n, int nrhs, const double *A, int lda,
                       const double *X, int ldx, const double *B, int ldb)
{
    if (n == 0 || nrhs == 0) return 0.0;
    long double rnorm = 0.0L, xnorm = 0.0L;
    for (int i = 0; i < n; ++i) {
        long double rrow = 0.0L, xrow = 0.0L;
        for (int c = 0; c < nrhs; ++c) {
            long double s = B_(i, c);
            for (int k = 0; k < n; ++k) s -= (long double)A_(i, k) * X[k + (size_t)c * ldx];
            rrow += fabsl(s);
            xrow += fabsl(X[i + (size_t)c * ldx]);
        }
        if (rrow > rnorm || isnan(rrow)) rnorm = rrow;
        if (xrow > xnorm || isnan(xrow)) xnorm = xrow;
    }
    const long double anorm = orc_dlange('I', n, n, A, lda);
    if (anorm == 0.0L || xnorm == 0.0L) return (rnorm == 0.0L) ? 0.0 : INFINITY;
    return (double)(rnorm / (n * anorm * xnorm) / orc_eps());
}

/* Cholesky factor check: ||A - L*L'||_F / (n * ||A||_F), divided by eps, over
 * the full symmetric matrix (lower: A = L L'; upper: A = U' U). */
double orc_chol_error(char uplo, int n, const double *A, int lda, const double *L, int ldl)
{
    if (n == 0) return 0.0;
    long double num = 0.0L, den = 0.0L;
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < n; ++i) {
            const int lo = (i > j) ? i : j;       /* element (lo, hi) of the stored triangle */
            const int hi = (i > j) ? j : i;
            long double s = 0.0L;
            for (int k = 0; k <= hi; ++k) {
                if (uplo == 'L') s += (long double)L[i + (size_t)k * ldl] * L[j + (size_t)k * ldl];
                else             s += (long double)L[k + (size_t)i * ldl] * L[k + (size_t)j * ldl];
            }
            long double a = (uplo == 'L') ? A_(lo, hi) : A_(hi, lo);
            num += (s - a) * (s - a);
            den += a * a;
        }
    }
    if (den == 0.0L) return (num == 0.0L) ? 0.0 : INFINITY;
    return (double)(sqrtl(num) / (sqrtl(den) * n) / orc_eps());
}

/* ---- batched drivers: the tester loop, OpenMP over matrices ------------ */
/* testing/testing_zgetrf_batched.cpp:188-212 */
void orc_batch_dgetrf(int n, double *A, int lda, int64_t stride_a, int *ipiv,
                      int *info, int batch, int nthreads)
{
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int s = 0; s < batch; ++s)
        orc_dgetrf(n, n, A + s * stride_a, lda, ipiv + (size_t)s * n, &info[s]);
}

/* testing/testing_zgesv_batched.cpp:159-186 (lapackf77_zgesv = getrf + getrs) */
void orc_batch_dgesv(int n, int nrhs, double *A, int lda, int64_t stride_a,
                     int *ipiv, double *B, int ldb, int64_t stride_b,
                     int *info, int batch, int nthreads)
{
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int s = 0; s < batch; ++s) {
        double *As = A + s * stride_a;
        int *ps = ipiv + (size_t)s * n;
        orc_dgetrf(n, n, As, lda, ps, &info[s]);
        if (info[s] == 0) orc_dgetrs('N', n, nrhs, As, lda, ps, B + s * stride_b, ldb);
    }
}

/* testing/testing_zposv_batched.cpp (lapackf77_zposv = potrf + potrs) */
void orc_batch_dposv(char uplo, int n, int nrhs, double *A, int lda, int64_t stride_a,
                     double *B, int ldb, int64_t stride_b, int *info, int batch,
                     int nthreads)
{
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int s = 0; s < batch; ++s) {
        double *As = A + s * stride_a;
        orc_dpotrf(uplo, n, As, lda, &info[s]);
        if (info[s] == 0) orc_dpotrs(uplo, n, nrhs, As, lda, B + s * stride_b, ldb);
    }
}

void orc_vbatch_dgetrf(const int *n, double *A, const int *lda, const int64_t *off,
                       int *ipiv, const int64_t *ipiv_off, int *info, int batch,
                       int nthreads)
{
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int s = 0; s < batch; ++s)
        orc_dgetrf(n[s], n[s], A + off[s], lda[s], ipiv + ipiv_off[s], &info[s]);
}

/* testing/testing_zgemm_vbatched.cpp:273-305 */
void orc_vbatch_dgemm(const int *m, const int *n, const int *k, double alpha,
                      const double *A, const int *lda, const int64_t *offa,
                      const double *B, const int *ldb, const int64_t *offb,
                      double beta, double *C, const int *ldc, const int64_t *offc,
                      int batch, int nthreads)
{
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int s = 0; s < batch; ++s)
        orc_dgemm('N', 'N', m[s], n[s], k[s], alpha, A + offa[s], lda[s],
                  B + offb[s], ldb[s], beta, C + offc[s], ldc[s]);
}
