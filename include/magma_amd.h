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

/* Batched dsgesv: FP32 LU with partial pivoting whose updates run on FP16
 * MFMA with a split (hi + lo) operand representation and FP32 accumulation,
 * after scaling each column by a power of two (largest entry in [0.5, 1));
 * FP64 residual refinement on the GPU.  ipiv receives the LAPACK-form pivots
 * of the FP32 factorization (of the FP64 one after a fallback). */
magma_int_t magma_dsgesv_batched(magma_int_t n, magma_int_t nrhs,
                                 double **dA_array, magma_int_t ldda, magma_int_t **dipiv_array,
                                 double **dB_array, magma_int_t lddb,
                                 double **dX_array, magma_int_t lddx,
                                 magma_int_t *dinfo_array, magma_int_t *diter_array,
                                 magma_int_t batchCount, magma_queue_t queue);

/* Batched dsposv: FP32 Cholesky (lower) of D A D with D = diag(2^-e_i),
 * a_ii 2^-2e_i in [0.5, 2), split-FP16 MFMA updates, FP64 refinement; batched
 * analogue of magma_dsposv_gpu (src/zcposv_gpu.cpp, include/magma_zc.h:68-77)
 * with the same argument meaning and the same iter codes as
 * magma_dsgesv_batched.  MagmaUpper reads the upper triangle (A = U'U).  On
 * fallback A holds the FP64 Cholesky factor in the triangle given by uplo and info[i] > 0 reports a matrix that is not
 * positive definite. */
magma_int_t magma_dsposv_batched(magma_uplo_t uplo, magma_int_t n, magma_int_t nrhs,
                                 double **dA_array, magma_int_t ldda,
                                 double **dB_array, magma_int_t lddb,
                                 double **dX_array, magma_int_t lddx,
                                 magma_int_t *dinfo_array, magma_int_t *diter_array,
                                 magma_int_t batchCount, magma_queue_t queue);

/* Variable-size dsgesv / dsposv: the same per-matrix semantics and argument
 * codes as the two above, with n, nrhs and the leading dimensions in device
 * arrays of batchCount+1 entries (the vbatched convention of
 * include/magma_zvbatched.h).  Orders above 512 take the FP64 path with
 * iter = -1; orders above 8192 return MAGMA_ERR_NOT_SUPPORTED.  The checked
 * vbatched calls read the checker's flags back, so they synchronize. */
magma_int_t magma_dsgesv_vbatched(magma_int_t *n, magma_int_t *nrhs,
                                  double **dA_array, magma_int_t *ldda, magma_int_t **dipiv_array,
                                  double **dB_array, magma_int_t *lddb,
                                  double **dX_array, magma_int_t *lddx,
                                  magma_int_t *dinfo_array, magma_int_t *diter_array,
                                  magma_int_t batchCount, magma_queue_t queue);
magma_int_t magma_dsposv_vbatched(magma_uplo_t uplo, magma_int_t *n, magma_int_t *nrhs,
                                  double **dA_array, magma_int_t *ldda,
                                  double **dB_array, magma_int_t *lddb,
                                  double **dX_array, magma_int_t *lddx,
                                  magma_int_t *dinfo_array, magma_int_t *diter_array,
                                  magma_int_t batchCount, magma_queue_t queue);

/* ---- autotuner --------------------------------------------------------------
 * Several kernel families have more than one implementation; which one runs
 * for a size class comes from a table (family, size class) -> variant that
 * replaces the reference's static tuning tables
 * (control/get_batched_crossover.cpp, control/batched_kernel_param.h,
 * control/get_batched_gemm_decision.cpp).  magma_amd_autotune times every
 * variant of every family on the queue's device with synthetic matrices of
 * each size class and keeps the fastest.  Untuned entries use variant 0.
 * Families and their variants:
 *   0 LU, 64 < n <= 512: 0 MFMA kernel, 1 the same at one workgroup per CU,
 *     2 VALU blocked kernel, 3 / 4 register-resident kernel at three / two
 *     workgroups per CU (size class 4, n <= 128, only);
 *   1 mixed-precision solver: five variants for n <= 256, two above;
 *   2 DGEMM: 0 32x32, 1 16x16, 2 48x48 tiles per wavefront, 3 64x64 tiles
 *     per workgroup staged through LDS.
 * magma_amd_tune_set returns -3 for a variant the size class does not have.
 * Size classes: 0..8 for max dimension <= 8, 16, 32, ..., 2048; 9 = larger
 * (up to 8192 for the factorizations) or unknown (vbatched calls without
 * tight bounds). */
#define MAGMA_AMD_TUNE_GETRF   1
#define MAGMA_AMD_TUNE_MIXED   2
#define MAGMA_AMD_TUNE_GEMM    4
#define MAGMA_AMD_TUNE_VERBOSE 256   /* print every timing on stderr */
magma_int_t magma_amd_autotune(magma_queue_t queue, magma_int_t flags);
/* variant of (family, size class), -1 when untuned */
magma_int_t magma_amd_tune_get(magma_int_t family, magma_int_t size_class);
/* variant -1 restores the default */
magma_int_t magma_amd_tune_set(magma_int_t family, magma_int_t size_class, magma_int_t variant);
magma_int_t magma_amd_tune_save(const char *path);
magma_int_t magma_amd_tune_load(const char *path);

/* ---- request server ------------------------------------------------------
 * Continuous batching of independent solves: requests are submitted from any
 * host thread, admitted into device memory under a fixed budget, batched
 * whenever one of the server's streams is free, and solved with the
 * mixed-precision kernels above (FP64 answers, FP64 fallback).  The client's
 * buffers are only read (dA, dB) or written (dX); the server works on its own
 * copy of the matrix.  A request with factor_key >= 0 keeps its factorization
 * resident after it completes; a later request with the same key, order and
 * kind reuses it (only the refinement runs).  Resident factorizations are
 * evicted least recently used first when memory is needed.  No reference
 * counterpart: the reference has no request loop. */
typedef enum {
    MAGMA_AMD_SOLVE_LU  = 0,    /* general matrix: LU with partial pivoting */
    MAGMA_AMD_SOLVE_SPD = 1     /* symmetric positive definite: Cholesky (lower part read) */
} magma_amd_solve_kind_t;

typedef struct {
    magma_int_t kind;                     /* magma_amd_solve_kind_t */
    magma_int_t n, nrhs;                  /* 1 <= n <= 512, nrhs >= 1 */
    const double *dA; magma_int_t ldda;   /* device, column-major */
    const double *dB; magma_int_t lddb;   /* device */
    double *dX; magma_int_t lddx;         /* device, receives the solution */
    long long factor_key;                 /* < 0: none; else keep / reuse the factorization under this key */
} magma_amd_request_t;

typedef struct {
    long long ticket;                     /* as returned by magma_amd_server_submit */
    magma_int_t info;                     /* > 0: singular / not positive definite (FP64 fallback) */
    magma_int_t iter;                     /* refinement sweeps (>= 0) or the dsgesv fallback code (< 0) */
    magma_int_t reused;                   /* 1 if a resident factorization served the request */
    double berr;                          /* ||b - A x||_inf / (||A||_inf ||x||_inf) of the mixed solution */
    double t_submit, t_start, t_done;     /* seconds on the server's steady clock */
} magma_amd_completion_t;

typedef struct {
    long long submitted, completed, batches, evictions;
    size_t budget_bytes, bytes_in_use, bytes_peak;
    magma_int_t resident, max_batch_seen;
} magma_amd_server_stats_t;

typedef struct magma_amd_server *magma_amd_server_t;

/* budget_bytes: device memory for matrix copies and factorizations (allocated
 * once); max_batch: most requests per launched batch. */
magma_int_t magma_amd_server_create(magma_device_t device, size_t budget_bytes, magma_int_t max_batch,
                                    magma_amd_server_t *server);
/* Returns the request's ticket (>= 0) or a negative argument / allocation code. */
long long magma_amd_server_submit(magma_amd_server_t server, const magma_amd_request_t *req);
/* Completed requests, without blocking; returns how many were written. */
magma_int_t magma_amd_server_poll(magma_amd_server_t server, magma_amd_completion_t *out, magma_int_t max_out);
/* Same, but waits up to timeout_s for at least one completion. */
magma_int_t magma_amd_server_wait(magma_amd_server_t server, magma_amd_completion_t *out, magma_int_t max_out,
                                  double timeout_s);
/* Waits until every submitted request has completed. */
magma_int_t magma_amd_server_drain(magma_amd_server_t server);
magma_int_t magma_amd_server_stats(magma_amd_server_t server, magma_amd_server_stats_t *stats);
/* Submits reqs[i] at arrival_s[i] seconds after the call (busy-waiting on the
 * clock) and returns when all have completed; out[i] is request i's record. */
magma_int_t magma_amd_server_run_trace(magma_amd_server_t server, const magma_amd_request_t *reqs,
                                       const double *arrival_s, magma_int_t count, magma_amd_completion_t *out);
void magma_amd_server_destroy(magma_amd_server_t server);

#ifdef __cplusplus
}
#endif
#endif
