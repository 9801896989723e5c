/*
 * magma_dvbatched.h -- variable-size batched ("vbatched") double-precision
 * entry points.  Per-matrix sizes and leading dimensions are DEVICE arrays of
 * length batchCount+1 (the reference's convention: the extra slot is scratch
 * the library may overwrite, magmablas/vbatched_check.cu:36-41,
 * magmablas/vbatched_aux.cu:504-538).
 *
 * magma_dpotrf_vbatched and magmablas_dgemm_vbatched mirror
 * include/magma_zvbatched.h:59-131.  magma_dgetrf_vbatched, magma_dgetrs_vbatched
 * and magma_dpotrs_vbatched have no counterpart in the reference (2.2.0 has no
 * vbatched LU or solves); they follow the same vbatched conventions and the
 * argument order of their fixed-size counterparts in magma_dbatched.h.
 */
#ifndef MAGMA_AMD_DVBATCHED_H
#define MAGMA_AMD_DVBATCHED_H

#include "magma_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* include/magma_zvbatched.h:59-70 */
magma_int_t magma_dpotrf_vbatched(magma_uplo_t uplo, magma_int_t *n, double **dA_array, magma_int_t *ldda,
                                  magma_int_t *info_array, magma_int_t batchCount, magma_queue_t queue);
magma_int_t magma_dpotrf_vbatched_max_nocheck(magma_uplo_t uplo, magma_int_t *n, double **dA_array,
                                              magma_int_t *ldda, magma_int_t *info_array,
                                              magma_int_t batchCount, magma_int_t max_n, magma_queue_t queue);

/* include/magma_zvbatched.h:90-131 */
void magmablas_dgemm_vbatched(magma_trans_t transA, magma_trans_t transB,
                              magma_int_t *m, magma_int_t *n, magma_int_t *k, double alpha,
                              double const *const *dA_array, magma_int_t *ldda,
                              double const *const *dB_array, magma_int_t *lddb, double beta,
                              double **dC_array, magma_int_t *lddc,
                              magma_int_t batchCount, magma_queue_t queue);
void magmablas_dgemm_vbatched_max_nocheck(magma_trans_t transA, magma_trans_t transB,
                                          magma_int_t *m, magma_int_t *n, magma_int_t *k, double alpha,
                                          double const *const *dA_array, magma_int_t *ldda,
                                          double const *const *dB_array, magma_int_t *lddb, double beta,
                                          double **dC_array, magma_int_t *lddc, magma_int_t batchCount,
                                          magma_int_t max_m, magma_int_t max_n, magma_int_t max_k,
                                          magma_queue_t queue);
void magmablas_dgemm_vbatched_max(magma_trans_t transA, magma_trans_t transB,
                                  magma_int_t *m, magma_int_t *n, magma_int_t *k, double alpha,
                                  double const *const *dA_array, magma_int_t *ldda,
                                  double const *const *dB_array, magma_int_t *lddb, double beta,
                                  double **dC_array, magma_int_t *lddc, magma_int_t batchCount,
                                  magma_int_t max_m, magma_int_t max_n, magma_int_t max_k,
                                  magma_queue_t queue);
void magmablas_dgemm_vbatched_nocheck(magma_trans_t transA, magma_trans_t transB,
                                      magma_int_t *m, magma_int_t *n, magma_int_t *k, double alpha,
                                      double const *const *dA_array, magma_int_t *ldda,
                                      double const *const *dB_array, magma_int_t *lddb, double beta,
                                      double **dC_array, magma_int_t *lddc,
                                      magma_int_t batchCount, magma_queue_t queue);

/* New (no reference counterpart): LU with partial pivoting of matrices of
 * different sizes; same semantics per matrix as magma_dgetrf_batched. */
magma_int_t magma_dgetrf_vbatched(magma_int_t *m, magma_int_t *n, double **dA_array, magma_int_t *ldda,
                                  magma_int_t **dipiv_array, magma_int_t *info_array,
                                  magma_int_t batchCount, magma_queue_t queue);
/* New: the same without argument checks and without any host synchronization
 * (graph-capturable); max_mn bounds max(m[i], n[i]). */
magma_int_t magma_dgetrf_vbatched_max_nocheck(magma_int_t *m, magma_int_t *n, magma_int_t max_mn,
                                              double **dA_array, magma_int_t *ldda,
                                              magma_int_t **dipiv_array, magma_int_t *info_array,
                                              magma_int_t batchCount, magma_queue_t queue);
/* No counterpart in the reference: magma_dgetrf_vbatched_max_nocheck that
 * allocates nothing (no hipMalloc, no queue workspace), for callers that
 * cannot allocate (HIP graph capture, serving loops, harnesses that forbid
 * it).  Its work list is a static device array holding at most 1048576
 * matrices; a larger batchCount returns -8.  Calls that overlap on one device
 * share that array and race. */
magma_int_t magma_dgetrf_vbatched_max_nocheck_noalloc(magma_int_t *m, magma_int_t *n, magma_int_t max_mn,
                                                      double **dA_array, magma_int_t *ldda,
                                                      magma_int_t **dipiv_array, magma_int_t *info_array,
                                                      magma_int_t batchCount, magma_queue_t queue);
/* New: solves with the getrf_vbatched / potrf_vbatched factors. */
magma_int_t magma_dgetrs_vbatched(magma_trans_t trans, magma_int_t *n, magma_int_t *nrhs,
                                  double **dA_array, magma_int_t *ldda, magma_int_t **dipiv_array,
                                  double **dB_array, magma_int_t *lddb,
                                  magma_int_t batchCount, magma_queue_t queue);
magma_int_t magma_dpotrs_vbatched(magma_uplo_t uplo, magma_int_t *n, magma_int_t *nrhs,
                                  double **dA_array, magma_int_t *ldda, double **dB_array, magma_int_t *lddb,
                                  magma_int_t batchCount, magma_queue_t queue);
/* New: the two solves without argument checks and without any host
 * synchronization (graph-capturable); max_n bounds every n[i]. */
magma_int_t magma_dgetrs_vbatched_max_nocheck(magma_trans_t trans, magma_int_t *n, magma_int_t *nrhs,
                                              double **dA_array, magma_int_t *ldda, magma_int_t **dipiv_array,
                                              double **dB_array, magma_int_t *lddb,
                                              magma_int_t batchCount, magma_int_t max_n, magma_queue_t queue);
magma_int_t magma_dpotrs_vbatched_max_nocheck(magma_uplo_t uplo, magma_int_t *n, magma_int_t *nrhs,
                                              double **dA_array, magma_int_t *ldda, double **dB_array,
                                              magma_int_t *lddb, magma_int_t batchCount, magma_int_t max_n,
                                              magma_queue_t queue);

#ifdef __cplusplus
}
#endif
#endif
