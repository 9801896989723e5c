/*
 * oracle.h -- CPU restatement of the reference's CPU path for MAGMA's batched
 * small-matrix factor-and-solve routines.
 *
 * THIS IS TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker.  The
 * product (libmagma_amd.so) never links, loads or calls it.
 *
 * The reference's testers (testing/testing_zgetrf_batched.cpp:188-212,
 * testing/testing_zgesv_batched.cpp:159-186, testing/testing_zpotrf_batched.cpp:125-148,
 * testing/testing_zgemm_vbatched.cpp:273-305) call LAPACK once per matrix inside
 * `#pragma omp parallel for schedule(dynamic)`.  LAPACK itself is external to the
 * reference (include/magma_zlapack.h:78-130); this file restates the netlib
 * algorithms (dgetrf/dgetrf2/dgetf2/idamax/dlaswp/dgetrs, dpotrf/dpotrf2/dpotrs,
 * reference-BLAS dtrsm/dgemm/dsyrk loop orders) in C.
 *
 * Rounding contract: every multiply-subtract of a factorization or solve is an
 * explicit fma() and the file is compiled with -ffp-contract=off, so the result
 * is the one a netlib LAPACK + reference BLAS build with FMA contraction gives.
 * In netlib's loop orders each matrix element receives its updates one at a
 * time, in increasing k for factorizations (decreasing k for the upper solve),
 * so the restatement fixes every rounding of the factors.
 *
 * Integer convention: ipiv is 1-based (LAPACK, MAGMA); info > 0 is the 1-based
 * column of the first failure (zero pivot / non-positive or NaN diagonal).
 */
#ifndef MAGMA_AMD_ORACLE_H
#define MAGMA_AMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- LAPACK-semantics kernels (one matrix) ---------------------------- */
int  orc_idamax(int n, const double *x);                       /* 0-based result */
void orc_dgetrf(int m, int n, double *A, int lda, int *ipiv, int *info);
void orc_dlaswp(int ncol, double *A, int lda, int k1, int k2, const int *ipiv, int incx);
void orc_dgetrs(char trans, int n, int nrhs, const double *A, int lda,
                const int *ipiv, double *B, int ldb);
void orc_dpotrf(char uplo, int n, double *A, int lda, int *info);
void orc_dpotrs(char uplo, int n, int nrhs, const double *A, int lda, double *B, int ldb);
void orc_dgetrf_nopiv(int m, int n, double *A, int lda, int *info);
void orc_rbt_butterflies(int n, double *u, double *v);
void orc_dgesv_rbt(int n, int nrhs, double *A, int lda, double *B, int ldb, int *info);
void orc_dgeqrf(int m, int n, double *A, int lda, double *tau);
void orc_dtrsm(char side, char uplo, char transa, char diag, int m, int n, double alpha,
               const double *A, int lda, double *B, int ldb);
void orc_dtrsv(char uplo, char trans, char diag, int n, const double *A, int lda, double *x, int incx);
void orc_dgemv(char trans, int m, int n, double alpha, const double *A, int lda,
               const double *x, int incx, double beta, double *y, int incy);
void orc_dgemm(char transa, char transb, int m, int n, int k, double alpha,
               const double *A, int lda, const double *B, int ldb,
               double beta, double *C, int ldc);

/* ---- tester error metrics (testing/testing_zgetrf_batched.cpp:30-70 etc.) */
double orc_eps(void);                                           /* dlamch('E') = 2^-53 */
double orc_dlange(char norm, int m, int n, const double *A, int lda);
double orc_lu_error(int m, int n, const double *A, int lda,
                    const double *LU, int ldlu, const int *ipiv);
double orc_solve_error(int n, int nrhs, const double *A, int lda,
                       const double *X, int ldx, const double *B, int ldb);
double orc_chol_error(char uplo, int n, const double *A, int lda, const double *L, int ldl);

/* ---- batched drivers: the reference CPU path (OpenMP, dynamic schedule) --- */
/* Matrices are stored back to back: matrix i at A + i*stride_a. */
void orc_batch_dgetrf(int n, double *A, int lda, int64_t stride_a, int *ipiv,
                      int *info, int batch, int nthreads);
void orc_batch_dgesv(int n, int nrhs, double *A, int lda, int64_t stride_a,
                     int *ipiv, double *B, int ldb, int64_t stride_b,
                     int *info, int batch, int nthreads);
void orc_batch_dposv(char uplo, int n, int nrhs, double *A, int lda, int64_t stride_a,
                     double *B, int ldb, int64_t stride_b, int *info, int batch,
                     int nthreads);
/* Variable sizes: matrix i has order n[i], leading dimension lda[i], offset off[i]. */
void orc_vbatch_dgetrf(const int *n, double *A, const int *lda, const int64_t *off,
                       int *ipiv, const int64_t *ipiv_off, int *info, int batch,
                       int nthreads);
void orc_vbatch_dgemm(const int *m, const int *n, const int *k, double alpha,
                      const double *A, const int *lda, const int64_t *offa,
                      const double *B, const int *ldb, const int64_t *offb,
                      double beta, double *C, const int *ldc, const int64_t *offc,
                      int batch, int nthreads);

/* ---- third-party LAPACK (OpenBLAS) loaded at run time -------------------
 * The reference's CPU path is whatever LAPACK MAGMA is linked with.  Where an
 * LP64 OpenBLAS with scipy_-prefixed symbols is present (the one bundled with
 * scipy), these run the same batched loops through it, single-threaded inside
 * LAPACK, for cross-checks and as the timed CPU baseline.  Return 0 on success. */
int  orc_lapack_open(const char *path);
const char *orc_lapack_config(void);
int  orc_lapack_batch_dgesv(int n, int nrhs, double *A, int lda, int64_t stride_a,
                            int *ipiv, double *B, int ldb, int64_t stride_b,
                            int *info, int batch, int nthreads);
int  orc_lapack_batch_dposv(char uplo, int n, int nrhs, double *A, int lda, int64_t stride_a,
                            double *B, int ldb, int64_t stride_b, int *info, int batch,
                            int nthreads);
int  orc_lapack_batch_dgetrf(int n, double *A, int lda, int64_t stride_a, int *ipiv,
                             int *info, int batch, int nthreads);

/* Wall-clock seconds (omp_get_wtime). */
double orc_wtime(void);
int    orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
