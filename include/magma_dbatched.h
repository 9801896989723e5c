/*
 * magma_dbatched.h -- fixed-size batched double-precision factor / solve /
 * GEMM entry points.  Each prototype is the reference's z-precision prototype
 * (include/magma_zbatched.h, cited per function) with magmaDoubleComplex ->
 * double, i.e. what tools/codegen.py generates as magma_dbatched.h.
 *
 * Conventions kept from the reference: all arrays device-resident (including
 * the pointer arrays); column-major matrices; ipiv 1-based; info_array zeroed
 * by the callee, info[i] > 0 a numerical failure of matrix i; return 0 or
 * -(position of the first bad argument) after magma_xerbla, or a MAGMA_ERR_*
 * code.  Every routine is enqueued on `queue` and the factor/solve routines
 * synchronize the queue before returning, as the reference's drivers do
 * (src/zgetrf_batched.cpp:354, src/zgetrs_batched.cpp:261,
 * src/zpotrf_batched.cpp:338, src/zpotrs_batched.cpp:237).
 */
#ifndef MAGMA_AMD_DBATCHED_H
#define MAGMA_AMD_DBATCHED_H

#include "magma_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* include/magma_zbatched.h:464-470 -- LU with partial pivoting, A = P*L*U */
magma_int_t magma_dgetrf_batched(magma_int_t m, magma_int_t n, double **dA_array, magma_int_t lda,
                                 magma_int_t **ipiv_array, magma_int_t *info_array,
                                 magma_int_t batchCount, magma_queue_t queue);

/* include/magma_zbatched.h:395-401 -- solve op(A) X = B with the getrf factors */
magma_int_t magma_dgetrs_batched(magma_trans_t trans, magma_int_t n, magma_int_t nrhs,
                                 double **dA_array, magma_int_t ldda, magma_int_t **dipiv_array,
                                 double **dB_array, magma_int_t lddb,
                                 magma_int_t batchCount, magma_queue_t queue);

/* include/magma_zbatched.h:617-623 -- getrf + getrs */
magma_int_t magma_dgesv_batched(magma_int_t n, magma_int_t nrhs, double **dA_array, magma_int_t ldda,
                                magma_int_t **dipiv_array, double **dB_array, magma_int_t lddb,
                                magma_int_t *dinfo_array, magma_int_t batchCount, magma_queue_t queue);

/* include/magma_zbatched.h:324-328 -- Cholesky A = L*L' (MagmaLower) or
 * A = U'*U (MagmaUpper, LAPACK's upper dpotrf; the reference's batched kernels
 * reject Upper with -1, magmablas/zpotf2_kernels.cu:153-156).  Upper reads and
 * writes only the upper triangle. */
magma_int_t magma_dpotrf_batched(magma_uplo_t uplo, magma_int_t n, double **dA_array, magma_int_t lda,
                                 magma_int_t *info_array, magma_int_t batchCount, magma_queue_t queue);

/* include/magma_zbatched.h:380-385 -- solve with the potrf factor (Lower or Upper) */
magma_int_t magma_dpotrs_batched(magma_uplo_t uplo, magma_int_t n, magma_int_t nrhs,
                                 double **dA_array, magma_int_t ldda, double **dB_array, magma_int_t lddb,
                                 magma_int_t batchCount, magma_queue_t queue);

/* include/magma_zbatched.h:387-393 -- potrf + potrs */
magma_int_t magma_dposv_batched(magma_uplo_t uplo, magma_int_t n, magma_int_t nrhs,
                                double **dA_array, magma_int_t ldda, double **dB_array, magma_int_t lddb,
                                magma_int_t *dinfo_array, magma_int_t batchCount, magma_queue_t queue);

/* include/magma_zbatched.h:585-606 -- LU without pivoting (m, n <= 2048) and
 * its solves; netlib dgetf2's operation order with the pivot fixed on the
 * diagonal.  info[i] = k > 0: the k-th pivot is exactly zero (the column is
 * left unscaled and the factorization continues). */
magma_int_t magma_dgetrf_nopiv_batched(magma_int_t m, magma_int_t n, double **dA_array, magma_int_t ldda,
                                       magma_int_t *info_array, magma_int_t batchCount, magma_queue_t queue);
magma_int_t magma_dgetrs_nopiv_batched(magma_trans_t trans, magma_int_t n, magma_int_t nrhs,
                                       double **dA_array, magma_int_t ldda, double **dB_array, magma_int_t lddb,
                                       magma_int_t *info_array, magma_int_t batchCount, magma_queue_t queue);
magma_int_t magma_dgesv_nopiv_batched(magma_int_t n, magma_int_t nrhs, double **dA_array, magma_int_t ldda,
                                      double **dB_array, magma_int_t lddb, magma_int_t *info_array,
                                      magma_int_t batchCount, magma_queue_t queue);

/* include/magma_zbatched.h:715-721 -- Householder QR (dgeqrf's reflectors:
 * R on and above the diagonal, the vectors below it, tau[i]); m, n <= 2048. */
magma_int_t magma_dgeqrf_batched(magma_int_t m, magma_int_t n, double **dA_array, magma_int_t lda,
                                 double **dtau_array, magma_int_t *info_array, magma_int_t batchCount,
                                 magma_queue_t queue);

/* include/magma_zbatched.h:609-614 -- solve by a depth-2 random butterfly
 * transformation (A <- U' A V, B <- U' B), LU without pivoting and X = V Y;
 * n <= 2048.  The butterfly weights are a fixed function of n.  Unlike the
 * reference, the caller's ldda is honoured, every right-hand side is
 * transformed, and the transform is exact for every n (the reference's is
 * consistent only for n a multiple of 4).  Synchronizes once to stage the
 * weights. */
magma_int_t magma_dgesv_rbt_batched(magma_int_t n, magma_int_t nrhs, double **dA_array, magma_int_t ldda,
                                    double **dB_array, magma_int_t lddb, magma_int_t *dinfo_array,
                                    magma_int_t batchCount, magma_queue_t queue);

/* include/magma_zbatched.h:473-479 -- inverse from the getrf factors (A and
 * ipiv as magma_dgetrf_batched leaves them) into invA; n <= 2048.  info[i] = k
 * > 0 when U(k,k) is exactly zero (the inverse is still formed). */
magma_int_t magma_dgetri_outofplace_batched(magma_int_t n, double **dA_array, magma_int_t ldda,
                                            magma_int_t **dipiv_array, double **dinvA_array, magma_int_t lddia,
                                            magma_int_t *info_array, magma_int_t batchCount, magma_queue_t queue);

/* include/magma_zbatched.h:202-208 -- B := alpha inv(op(A)) B (Left) or
 * alpha B inv(op(A)) (Right), A triangular; the operation order of the
 * reference BLAS dtrsm, so results are those of LAPACK's CPU path bit for bit. */
void magmablas_dtrsm_batched(magma_side_t side, magma_uplo_t uplo, magma_trans_t transA, magma_diag_t diag,
                             magma_int_t m, magma_int_t n, double alpha,
                             double **dA_array, magma_int_t ldda, double **dB_array, magma_int_t lddb,
                             magma_int_t batchCount, magma_queue_t queue);
/* include/magma_zbatched.h:237-242 -- b := inv(op(A)) b in dtrsv's operation
 * order; n <= 4095, incb >= 1 (the reference's n / lda error codes, -5 / -8). */
void magmablas_dtrsv_batched(magma_uplo_t uplo, magma_trans_t trans, magma_diag_t diag, magma_int_t n,
                             double **A_array, magma_int_t lda, double **b_array, magma_int_t incb,
                             magma_int_t batchCount, magma_queue_t queue);
/* include/magma_zbatched.h:705-712 -- y := alpha op(A) x + beta y in dgemv's
 * operation order; incx, incy >= 1. */
void magmablas_dgemv_batched(magma_trans_t trans, magma_int_t m, magma_int_t n, double alpha,
                             double **dA_array, magma_int_t ldda, double **dx_array, magma_int_t incx,
                             double beta, double **dy_array, magma_int_t incy,
                             magma_int_t batchCount, magma_queue_t queue);

/* include/magma_zbatched.h:88-108 -- C = alpha op(A) op(B) + beta C, fixed sizes */
void magma_dgemm_batched(magma_trans_t transA, magma_trans_t transB,
                         magma_int_t m, magma_int_t n, magma_int_t k, double alpha,
                         double const *const *dA_array, magma_int_t ldda,
                         double const *const *dB_array, magma_int_t lddb, double beta,
                         double **dC_array, magma_int_t lddc, magma_int_t batchCount, magma_queue_t queue);
void magmablas_dgemm_batched(magma_trans_t transA, magma_trans_t transB,
                             magma_int_t m, magma_int_t n, magma_int_t k, double alpha,
                             double const *const *dA_array, magma_int_t ldda,
                             double const *const *dB_array, magma_int_t lddb, double beta,
                             double **dC_array, magma_int_t lddc, magma_int_t batchCount, magma_queue_t queue);

/* include/magma_zbatched.h:28-41 -- pointer-array helpers
 * output_array[i] = input + row + column*lda + i*batch_offset */
void magma_dset_pointer(double **output_array, double *input, magma_int_t lda,
                        magma_int_t row, magma_int_t column, magma_int_t batch_offset,
                        magma_int_t batchCount, magma_queue_t queue);
/* output_array[i] = input_array[i] + row + column*lda */
void magma_ddisplace_pointers(double **output_array, double **input_array, magma_int_t lda,
                              magma_int_t row, magma_int_t column,
                              magma_int_t batchCount, magma_queue_t queue);
/* include/magma_batched.h:47-53 -- output_array[i] = input + row + column*lda + i*batchSize */
void magma_iset_pointer(magma_int_t **output_array, magma_int_t *input, magma_int_t lda,
                        magma_int_t row, magma_int_t column, magma_int_t batchSize,
                        magma_int_t batchCount, magma_queue_t queue);

#ifdef __cplusplus
}
#endif
#endif
