// solve.hip -- triangular solves with LU / Cholesky factors (getrs, potrs),
// one 256-thread workgroup per matrix, right-hand sides staged in LDS in
// chunks of up to kChunk columns.
//
// Replaces the reference's getrs/potrs chains (src/zgetrs_batched.cpp:86-273,
// src/zpotrs_batched.cpp:70-249: laswp_rowserial + trsv, or trtri_diag +
// trsm_outofplace GEMMs with a 128x128 dinvA workspace per matrix) with one
// launch and no workspace.  Sweeps are column-oriented (axpy form, one barrier
// per step): for the no-transpose L and U solves this is the reference BLAS
// dtrsm loop order, so results are bit-identical to LAPACK's dgetrs /
// dpotrs forward sweep; the transposed sweeps associate their sums in the
// opposite order to dtrsm's dot-product form and agree to rounding.
#include <algorithm>

#include "kernels.h"

namespace magma_amd {
namespace {

constexpr int kThreads = 256;
constexpr int kChunk = 8;

__device__ __forceinline__ int pick(const int *arr, int fixed, int id) { return arr ? arr[id] : fixed; }

// One triangular sweep over X (n x nc, LDS, leading dimension ldx) with factor T.
//   lower/upper: which triangle of T is used; trans: op(T) = T';
//   unit: implicit unit diagonal.  op(T) lower -> forward sweep, upper -> backward.
__device__ void sweep(const double *T, int ldt, int n, double *X, int ldx, int nc,
                      bool lower, bool trans, bool unit)
{
    const int tid = threadIdx.x;
    const bool forward = (lower != trans);
    if (n <= kThreads) {
        // one row per thread: the next step's entries of T (its column, or row
        // when transposed, and its diagonal) are loaded while the current step
        // runs, so each step does not wait on a global load
        auto entry = [&](int k, int i) { return trans ? T[k + (size_t)i * ldt] : T[i + (size_t)k * ldt]; };
        auto row_of = [&](int k) { return (forward ? k + 1 : 0) + tid; };
        auto row_end = [&](int k) { return forward ? n : k; };
        int k = forward ? 0 : n - 1;
        double tc = 0.0, dc = 1.0;
        if (row_of(k) < row_end(k)) tc = entry(k, row_of(k));
        if (!unit && tid < nc) dc = T[k + (size_t)k * ldt];
        for (int s = 0; s < n; ++s) {
            k = forward ? s : n - 1 - s;
            double tn = 0.0, dn = 1.0;
            if (s + 1 < n) {
                const int k1 = forward ? s + 1 : n - 2 - s;
                if (row_of(k1) < row_end(k1)) tn = entry(k1, row_of(k1));
                if (!unit && tid < nc) dn = T[k1 + (size_t)k1 * ldt];
            }
            if (!unit) {
                if (tid < nc) {
                    double *xk = X + (size_t)tid * ldx + k;
                    // no-transpose sweeps skip a zero entry like reference dtrsm;
                    // its transposed (dot-form) sweeps always divide
                    if (trans || *xk != 0.0) *xk = *xk / dc;
                }
                __syncthreads();
            }
            const int i = row_of(k);
            if (i < row_end(k)) {
                for (int c = 0; c < nc; ++c) {
                    const double xk = X[(size_t)c * ldx + k];
                    if (!trans && xk == 0.0) continue;      // reference dtrsm's skip (x_k stays 0 after division)
                    X[(size_t)c * ldx + i] = fma(-xk, tc, X[(size_t)c * ldx + i]);
                }
            }
            __syncthreads();
            tc = tn;
            dc = dn;
        }
        return;
    }
    for (int s = 0; s < n; ++s) {
        const int k = forward ? s : n - 1 - s;
        if (!unit) {
            if (tid < nc) {
                double *xk = X + (size_t)tid * ldx + k;
                const double d = T[k + (size_t)k * ldt];
                // no-transpose sweeps skip a zero entry like reference dtrsm;
                // its transposed (dot-form) sweeps always divide
                if (trans || *xk != 0.0) *xk = *xk / d;
            }
            __syncthreads();
        }
        // rows still to update: i > k (forward) or i < k (backward)
        const int lo = forward ? k + 1 : 0;
        const int hi = forward ? n : k;
        for (int i = lo + tid; i < hi; i += kThreads) {
            // op(T)(i,k): T(i,k) without transpose, T(k,i) with
            const double t = trans ? T[k + (size_t)i * ldt] : T[i + (size_t)k * ldt];
            for (int c = 0; c < nc; ++c) {
                const double xk = X[(size_t)c * ldx + k];
                if (!trans && xk == 0.0) continue;      // reference dtrsm's skip (x_k stays 0 after division)
                X[(size_t)c * ldx + i] = fma(-xk, t, X[(size_t)c * ldx + i]);
            }
        }
        __syncthreads();
    }
}

// mode: 'N' getrs no-trans, 'T' getrs trans, 'L' potrs lower, 'U' potrs upper.
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) solve_kernel(MatSet s, char mode, int chunk)
{
    extern __shared__ double X[];
    const int cnt = item_count(s);
    const int *idx = item_index(s);
    for (int item = blockIdx.x; item < cnt; item += gridDim.x) {
        const int id = idx ? idx[item] : item;
        const int n = pick(s.n_arr, s.n, id);
        const int lda = pick(s.lda_arr, s.lda, id);
        const int nrhs = pick(s.nrhs_arr, s.nrhs, id);
        const int ldb = pick(s.ldb_arr, s.ldb, id);
        const double *A = s.A[id];
        double *B = s.B[id];
        const int *ipiv = (mode == 'N' || mode == 'T') ? s.ipiv[id] : nullptr;
        const int tid = threadIdx.x;
        const int ldx = n;
        if (n <= 0) continue;
        // after a fused factorization: no solve with a singular factor (dgesv/dposv)
        if (s.info && s.info[id] != 0) continue;

        // the interchanges staged in LDS (after the right-hand-side chunk), so
        // the serial dlaswp chain below waits on LDS, not on a global load per step
        int *piv = reinterpret_cast<int *>(X + (size_t)n * chunk);
        if (ipiv) {
            for (int k = tid; k < n; k += kThreads) piv[k] = ipiv[k];
        }
        for (int c0 = 0; c0 < nrhs; c0 += chunk) {
            const int nc = min(chunk, nrhs - c0);
            for (int e = tid; e < n * nc; e += kThreads) {
                const int i = e % n, c = e / n;
                X[(size_t)c * ldx + i] = B[i + (size_t)(c0 + c) * ldb];
            }
            __syncthreads();
            if (mode == 'N') {
                // dlaswp forward: interchanges in order, one thread per column
                if (tid < nc)
                    for (int k = 0; k < n; ++k) {
                        const int p = piv[k] - 1;
                        if (p != k && p >= 0 && p < n) {
                            double *x = X + (size_t)tid * ldx;
                            const double t = x[k];
                            x[k] = x[p];
                            x[p] = t;
                        }
                    }
                __syncthreads();
                sweep(A, lda, n, X, ldx, nc, true, false, true);
                sweep(A, lda, n, X, ldx, nc, false, false, false);
            } else if (mode == 'T') {
                sweep(A, lda, n, X, ldx, nc, false, true, false);
                sweep(A, lda, n, X, ldx, nc, true, true, true);
                if (tid < nc)
                    for (int k = n - 1; k >= 0; --k) {
                        const int p = piv[k] - 1;
                        if (p != k && p >= 0 && p < n) {
                            double *x = X + (size_t)tid * ldx;
                            const double t = x[k];
                            x[k] = x[p];
                            x[p] = t;
                        }
                    }
                __syncthreads();
            } else if (mode == 'L') {
                sweep(A, lda, n, X, ldx, nc, true, false, false);
                sweep(A, lda, n, X, ldx, nc, true, true, false);
            } else {
                sweep(A, lda, n, X, ldx, nc, false, true, false);
                sweep(A, lda, n, X, ldx, nc, false, false, false);
            }
            for (int e = tid; e < n * nc; e += kThreads) {
                const int i = e % n, c = e / n;
                B[i + (size_t)(c0 + c) * ldb] = X[(size_t)c * ldx + i];
            }
            __syncthreads();
        }
        __syncthreads();
    }
}

}  // namespace

static magma_int_t launch_solve(const MatSet &s, char mode, int max_n, hipStream_t st)
{
    if (s.count <= 0 || max_n <= 0) return MAGMA_SUCCESS;
    // right-hand-side columns per pass: as many as fit 64 KB of LDS, at most kChunk
    const int chunk = std::min<long>(kChunk, (64 * 1024) / (sizeof(double) * (size_t)max_n));
    if (chunk < 1) return MAGMA_ERR_NOT_SUPPORTED;
    const size_t lds = sizeof(double) * (size_t)max_n * chunk + sizeof(int) * (size_t)max_n;
    const int grid = s.dev_count ? persistent_grid(s.count, 4) : s.count;
    solve_kernel<<<grid, kThreads, lds, st>>>(s, mode, chunk);
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

magma_int_t launch_getrs(const MatSet &s, char trans, int max_n, int max_nrhs, hipStream_t st)
{
    (void)max_nrhs;
    return launch_solve(s, trans == 'N' ? 'N' : 'T', max_n, st);
}

magma_int_t launch_potrs(const MatSet &s, char uplo, int max_n, int max_nrhs, hipStream_t st)
{
    (void)max_nrhs;
    return launch_solve(s, uplo == 'L' ? 'L' : 'U', max_n, st);
}

}  // namespace magma_amd
