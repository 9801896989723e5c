/*
 * lapack_dl.c -- the reference's CPU path through a third-party LAPACK
 * (OpenBLAS, LP64, scipy_-prefixed symbols) loaded with dlopen.
 * TEST INFRASTRUCTURE ONLY (cross-checks and the timed CPU baseline).
 *
 * The batched loops restate the reference testers: one LAPACK call per matrix
 * under `#pragma omp parallel for schedule(dynamic)` with LAPACK itself kept
 * single-threaded (testing/testing_zgetrf_batched.cpp:188-212,
 * testing/testing_zgesv_batched.cpp:159-186; thread control as in
 * control/magma_threadsetting.cpp:121-175).
 */
#include "oracle.h"

#include <dlfcn.h>
#include <stddef.h>

typedef void (*getrf_t)(const int *, const int *, double *, const int *, int *, int *);
typedef void (*getrs_t)(const char *, const int *, const int *, const double *, const int *,
                        const int *, double *, const int *, int *, size_t);
typedef void (*potrf_t)(const char *, const int *, double *, const int *, int *, size_t);
typedef void (*potrs_t)(const char *, const int *, const int *, const double *, const int *,
                        double *, const int *, int *, size_t);
typedef void (*setthreads_t)(int);
typedef char *(*config_t)(void);

static getrf_t p_getrf;
static getrs_t p_getrs;
static potrf_t p_potrf;
static potrs_t p_potrs;
static config_t p_config;

int orc_lapack_open(const char *path)
{
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1;
    p_getrf = (getrf_t)dlsym(h, "scipy_dgetrf_");
    p_getrs = (getrs_t)dlsym(h, "scipy_dgetrs_");
    p_potrf = (potrf_t)dlsym(h, "scipy_dpotrf_");
    p_potrs = (potrs_t)dlsym(h, "scipy_dpotrs_");
    p_config = (config_t)dlsym(h, "scipy_openblas_get_config");
    setthreads_t st = (setthreads_t)dlsym(h, "scipy_openblas_set_num_threads");
    if (!p_getrf || !p_getrs || !p_potrf || !p_potrs) return -2;
    if (st) st(1);          /* LAPACK single-threaded; parallelism is over the batch */
    return 0;
}

const char *orc_lapack_config(void) { return p_config ? p_config() : "unknown"; }

int orc_lapack_batch_dgetrf(int n, double *A, int lda, int64_t stride_a, int *ipiv,
                            int *info, int batch, int nthreads)
{
    if (!p_getrf) return -1;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int s = 0; s < batch; ++s)
        p_getrf(&n, &n, A + s * stride_a, &lda, ipiv + (size_t)s * n, &info[s]);
    return 0;
}

int orc_lapack_batch_dgesv(int n, int nrhs, double *A, int lda, int64_t stride_a,
                           int *ipiv, double *B, int ldb, int64_t stride_b,
                           int *info, int batch, int nthreads)
{
    if (!p_getrf) return -1;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int s = 0; s < batch; ++s) {
        double *As = A + s * stride_a;
        int *ps = ipiv + (size_t)s * n;
        int inf2 = 0;
        p_getrf(&n, &n, As, &lda, ps, &info[s]);
        if (info[s] == 0) p_getrs("N", &n, &nrhs, As, &lda, ps, B + s * stride_b, &ldb, &inf2, 1);
    }
    return 0;
}

int orc_lapack_batch_dposv(char uplo, int n, int nrhs, double *A, int lda, int64_t stride_a,
                           double *B, int ldb, int64_t stride_b, int *info, int batch,
                           int nthreads)
{
    if (!p_potrf) return -1;
    const char u[2] = {uplo, 0};
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int s = 0; s < batch; ++s) {
        double *As = A + s * stride_a;
        int inf2 = 0;
        p_potrf(u, &n, As, &lda, &info[s], 1);
        if (info[s] == 0) p_potrs(u, &n, &nrhs, As, &lda, B + s * stride_b, &ldb, &inf2, 1);
    }
    return 0;
}
