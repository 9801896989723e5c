// mixed.h -- launch interface of the mixed-precision solver kernels (FP32
// factorization, FP64 refinement) shared by the dsgesv / dsposv entry points
// and the request server.
#pragma once

#include "kernels.h"

namespace magma_amd {

// Largest order the mixed kernels take.
constexpr int kMixedMaxN = 512;

// Per-matrix factor storage, in floats: the n x n factor (ld n, rows in the
// input's original order), then n ints of scaling exponents, then n ints of
// pivot order (original row of pivot position i), then (LU) the diagonal
// 32 x 32 block of every 32-column chunk with its rows in pivot order,
// [chunk][i][k] = factor(p[32 chunk + i], 32 chunk + k), so that the block
// solves read it as one contiguous 4 KB piece instead of 32 scattered rows.
__host__ __device__ inline size_t mixed_diag_offset(int n) { return ((size_t)n * n + 2 * (size_t)n + 3) & ~(size_t)3; }
__host__ __device__ inline size_t mixed_factor_floats(int n)
{
    return mixed_diag_offset(n) + 1024 * (size_t)((n + 31) / 32);
}

struct MixedSet {
    int n = 0, nrhs = 0, lda = 0, ldb = 0, ldx = 0;       // fixed sizes
    const int *n_arr = nullptr, *nrhs_arr = nullptr;       // per matrix (nullable)
    const int *lda_arr = nullptr, *ldb_arr = nullptr, *ldx_arr = nullptr;
    char kind = 'L';                  // 'L' LU with partial pivoting, 'C' Cholesky (lower)
    const char *kind_arr = nullptr;   // per matrix (nullable)
    double **A = nullptr, **B = nullptr, **X = nullptr;
    float **F = nullptr;              // per-matrix factor storage (nullable)
    float *Wslot = nullptr;           // else one slot of mixed_factor_floats(max n) per workgroup
    size_t slot_floats = 0;
    int **ipiv = nullptr;             // LAPACK-form pivots of the FP32 factorization (nullable)
    int *info = nullptr, *iter = nullptr;
    double *berr = nullptr;           // final ||r||/(||A|| ||x||) per matrix (nullable)
    const int *reuse = nullptr;       // per matrix: nonzero -> F already holds the factorization
    int *fb_list = nullptr, *fb_count = nullptr;       // LU matrices that fall back to FP64
    int *fbc_list = nullptr, *fbc_count = nullptr;     // Cholesky matrices that fall back
    const int *index = nullptr;       // matrices to process (nullable: 0..count-1)
    int count = 0;
};

// Launch plan for `count` matrices of order at most max_n with at most
// max_nrhs right-hand sides each (0: not known on the host): the kernel
// variant, read from the tuning table once, and its persistent grid, so the
// factor slots sized from the plan (one per workgroup) are those the launch
// uses even if the table changes in between.
struct MixedPlan {
    int variant = 0;
    int grid = 0;
};
MixedPlan mixed_plan(int count, int max_n, int max_nrhs);
// Device workspace of a plan: one factor slot per workgroup when `slots`
// (the set has no per-matrix factor storage); the view sets g.Wslot and
// g.slot_floats.
size_t mixed_workspace_bytes(const MixedPlan &p, int max_n, bool slots);
void mixed_workspace_view(MixedSet &g, const MixedPlan &p, int max_n, bool slots, void *ws);
magma_int_t launch_mixed(const MixedSet &g, int max_n, const MixedPlan &plan, hipStream_t st);

// FP64 fallback for the matrices the mixed kernel listed: X <- B, then the
// plain factorization and solve in place (A receives the FP64 factors).
// `lists` must be the device lists of a completed-or-enqueued launch_mixed.
magma_int_t launch_mixed_fallback(const MixedSet &g, int max_n, int max_nrhs, hipStream_t st);

// X <- B for the matrices list[0 .. *count) (device list and count).
magma_int_t launch_copy_rhs(const MixedSet &g, const int *list, const int *count, hipStream_t st);
// *a = 0 and *b = 0 on the stream (either may be null).
magma_int_t launch_zero_counts(int *a, int *b, hipStream_t st);

}  // namespace magma_amd
