/*
 * magma_amd.h -- extensions of this library beyond the reference's API.
 *
 * Mixed-precision batched solve.  The reference has mixed precision only for a
 * single matrix (magma_dsgesv_gpu, src/zcgesv_gpu.cpp:143-365; prototype
 * include/magma_zc.h:34-44); these are its batched analogues with the same
 * argument meaning: A and B are not modified when refinement succeeds, X
 * receives the FP64 solution, iter[i] > 0 is the number of refinement sweeps
 * matrix i took, iter[i] < 0 means matrix i fell back to the plain FP64
 * factorization (then A holds its FP64 LU factors and ipiv its pivots):
 *   -(MAGMA_MIXED_ITERMAX+1)  not converged in MAGMA_MIXED_ITERMAX sweeps
 *   -3                        the low-precision factorization hit a zero pivot
 *   -2                        the matrix could not be scaled into range
 * info[i] > 0 only if the FP64 fallback factorization is singular.
 * Convergence test per right-hand side (src/zcgesv_gpu.cpp:266-279):
 *   ||r||_inf <= ||x||_inf * ||A||_inf * eps64 * sqrt(n).
 */
#ifndef MAGMA_AMD_EXT_H
#define MAGMA_AMD_EXT_H

#include "magma_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MAGMA_MIXED_ITERMAX 10

/* Batched dsgesv: FP32 LU whose trailing updates run on FP16 MFMA with a
 * split (hi + lo) operand representation and FP32 accumulation, after scaling
 * each matrix by 1/max|a_ij|; FP64 residual refinement on the GPU. */
magma_int_t magma_dsgesv_batched(magma_int_t n, magma_int_t nrhs,
                                 double **dA_array, magma_int_t ldda, magma_int_t **dipiv_array,
                                 double **dB_array, magma_int_t lddb,
                                 double **dX_array, magma_int_t lddx,
                                 magma_int_t *dinfo_array, magma_int_t *diter_array,
                                 magma_int_t batchCount, magma_queue_t queue);

#ifdef __cplusplus
}
#endif
#endif
