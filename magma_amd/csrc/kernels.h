// kernels.h -- launch interfaces between the C-ABI drivers and the HIP kernels.
#pragma once

#include "internal.h"

namespace magma_amd {

// Pointers in the global address space.  A device pointer read from a pointer
// array through one of these and converted back to a generic pointer lets the
// compiler emit global rather than flat accesses through it (a flat access
// also counts against the LDS wait counter, so LDS waits would drain it).
typedef double __attribute__((address_space(1))) *GDoublePtr;
typedef float __attribute__((address_space(1))) *GFloatPtr;
typedef int __attribute__((address_space(1))) *GIntPtr;

// Describes a set of matrices for one launch.  Sizes either fixed (arrays
// null) or per matrix (device arrays indexed by matrix id).  `index`, when not
// null, lists the matrix ids this launch processes (a bucket of a vbatched
// call); otherwise ids are 0..count-1.
struct MatSet {
    int m = 0, n = 0, lda = 0;                      // fixed sizes
    const int *m_arr = nullptr, *n_arr = nullptr, *lda_arr = nullptr;
    double **A = nullptr;
    int **ipiv = nullptr;
    int *info = nullptr;
    // right-hand sides (solves)
    int nrhs = 0, ldb = 0;
    const int *nrhs_arr = nullptr, *ldb_arr = nullptr;
    double **B = nullptr;
    const int *index = nullptr;
    // Device-resident work list (vbatched buckets): when dev_count is set the
    // number of items is *dev_count (count is then only an upper bound used
    // to size the grid) and the id list starts at index + *dev_offset.
    const int *dev_count = nullptr;
    const int *dev_offset = nullptr;
    int count = 0;
};

#ifdef __HIPCC__
__device__ __forceinline__ int item_count(const MatSet &s) { return s.dev_count ? *s.dev_count : s.count; }
__device__ __forceinline__ const int *item_index(const MatSet &s)
{
    return s.index ? s.index + (s.dev_offset ? *s.dev_offset : 0) : nullptr;
}
#endif

// Grid size for a persistent launch of `items` work units of which at most
// `per_cu` are resident per compute unit.
int persistent_grid(long items, int per_cu);

// LU with partial pivoting, LAPACK semantics (optionally followed by the
// no-transpose solve of the right-hand sides in `B`).  Chooses the kernel
// family from max_m / max_n.  `max_mn` bounds min(m,n) over the set.
magma_int_t launch_getrf(const MatSet &s, int max_m, int max_n, bool fused_solve, hipStream_t st);
// Cholesky, lower, LAPACK semantics (optionally fused with the solve).
magma_int_t launch_potrf(const MatSet &s, int max_n, bool fused_solve, hipStream_t st);
// Solves with existing factors.  trans: 'N' or 'T' (getrs); uplo 'L'/'U' (potrs).
magma_int_t launch_getrs(const MatSet &s, char trans, int max_n, int max_nrhs, hipStream_t st);
magma_int_t launch_potrs(const MatSet &s, char uplo, int max_n, int max_nrhs, hipStream_t st);

// Swaps A(i, j) and A(j, i) for every i > j in place (0 < n <= lda; other
// matrices are skipped).  Two swaps around the lower-triangular Cholesky
// kernels factor an upper-stored matrix with LAPACK's upper dpotrf arithmetic.
magma_int_t launch_swap_triangles(const MatSet &s, hipStream_t st);

// GEMM C = alpha op(A) op(B) + beta C for a set of (possibly different) sizes.
struct GemmSet {
    char ta = 'N', tb = 'N';
    int m = 0, n = 0, k = 0, lda = 0, ldb = 0, ldc = 0;
    const int *m_arr = nullptr, *n_arr = nullptr, *k_arr = nullptr;
    const int *lda_arr = nullptr, *ldb_arr = nullptr, *ldc_arr = nullptr;
    double alpha = 1.0, beta = 0.0;
    const double *const *A = nullptr;
    const double *const *B = nullptr;
    double **C = nullptr;
    const int *index = nullptr;
    int count = 0;
};
magma_int_t launch_gemm(const GemmSet &g, int max_m, int max_n, int max_k, hipStream_t st);

}  // namespace magma_amd
