// blas_batched.hip -- public batched triangular solves and matrix-vector
// products: magmablas_dtrsv_batched, magmablas_dtrsm_batched and
// magmablas_dgemv_batched (include/magma_zbatched.h:202-258,705-712).
//
// Every entry is computed with the operation sequence of the reference BLAS
// routine (dtrsv.f, dtrsm.f, dgemv.f; the testers compare against them,
// testing/testing_ztrsv_batched.cpp:192, testing/testing_ztrsm_batched.cpp:234,
// testing/testing_zgemv_batched.cpp:155), each update one fma, so results are
// bitwise those of the CPU restatement in the oracle.  The reference instead
// inverts diagonal blocks and multiplies (magmablas/ztrsm_batched.cpp:168-480,
// magmablas/ztrsv_batched.cu:70-296), which rounds differently; LAPACK's
// order is the target here.
//
//   trsv  one workgroup per matrix, x in LDS.  Each of the four cases is a
//         sequence of pivots (ascending or descending) whose final value
//         updates every remaining entry in parallel.  For the transposed
//         cases this "right-looking" order adds the terms of each dot product
//         in exactly dtrsv.f's order (ascending for Upper, descending for
//         Lower), so no case needs a serial dot.  The pivots' columns (rows,
//         transposed) of A are staged in LDS a panel at a time with coalesced
//         loads.
//   trsm  one lane per independent vector (a column of B for side Left, a
//         row for Right), each lane running dtrsm.f's loops for its vector,
//         held in LDS when it is at most 128 long; matrices with fewer than
//         64 vectors share a wavefront.  A's entries are the same for every
//         lane of a matrix, so they are broadcast loads.  Two cases (Left/Trans/Lower and Right/NoTrans/
//         Lower) add dot products in the order opposite to the order their
//         terms become final, which rules out a parallel right-looking form
//         for bit-exact results.
//   gemv  NoTrans: one thread per entry of y, columns in order (coalesced A
//         rows).  Trans: one thread per entry of y, a 64-column tile of A
//         staged through LDS 32 rows at a time, each thread summing its
//         column in row order.
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>

#include "kernels.h"

namespace magma_amd {

magma_int_t zero_info(int *info, int count, hipStream_t st);

namespace {

// Largest order the batched trsv takes (x and one pivot of A in 64 KB of LDS).
#define MAGMA_AMD_TRSV_MAX_N 4095

constexpr int kTrsvThreads = 256;
constexpr int kTrsmStagedMax = 128;        // vectors up to this length live in LDS (64 KB per wavefront)
constexpr size_t kTrsvLds = 64 * 1024;     // x plus the staged panel of A

// NT threads per matrix: one wavefront for short vectors (its barriers are
// nearly free), four otherwise.
template <bool TRANS, bool UPPER, int NT>
__global__ void __launch_bounds__(NT) trsv_kernel(int n, double **A_array, int lda, double **x_array,
                                                  int incx, bool unit, int pb, int count)
{
    extern __shared__ double lds[];
    double *x = lds;                 // n
    double *pan = lds + n;           // pb x n: pan[c * n + i]
    constexpr bool asc = (TRANS == UPPER);
    const int tid = threadIdx.x;
    for (int item = blockIdx.x; item < count; item += gridDim.x) {
        const double *a = A_array[item];
        double *xg = x_array[item];
        for (int i = tid; i < n; i += NT) x[i] = xg[(size_t)i * incx];
        for (int s0 = 0; s0 < n; s0 += pb) {
            const int np = min(pb, n - s0);
            __syncthreads();
            // stage the panel: pivot p_c = step s0 + c; NoTrans column p_c,
            // Trans row p_c, all n entries (c fastest for Trans, so each
            // column of A is read as one contiguous run of pivots' rows)
            constexpr int W = NT / 64;
            const int lane = tid & 63, wv = tid >> 6;
            if (!TRANS) {
#pragma unroll 8
                for (int c = wv; c < np; c += W) {
                    const int p = asc ? s0 + c : n - 1 - (s0 + c);
                    for (int i = lane; i < n; i += 64) pan[c * n + i] = a[i + (size_t)p * lda];
                }
            } else {
#pragma unroll 8
                for (int j = wv; j < n; j += W)
                    for (int c = lane; c < np; c += 64) {
                        const int p = asc ? s0 + c : n - 1 - (s0 + c);
                        pan[c * n + j] = a[p + (size_t)j * lda];
                    }
            }
            __syncthreads();
#pragma unroll 1
            for (int c = 0; c < np; ++c) {
                const int p = asc ? s0 + c : n - 1 - (s0 + c);
                const double *col = pan + c * n;
                const double xr = x[p];
                bool skip = false;
                double xp;
                if (!TRANS) {
                    // dtrsv.f NoTrans: the zero skip covers the division
                    skip = (xr == 0.0);
                    xp = (skip || unit) ? xr : xr / col[p];
                } else {
                    xp = unit ? xr : xr / col[p];
                }
                const int lo = asc ? p + 1 : 0, hi = asc ? n : p;
                if (!skip)
                    for (int i = lo + tid; i < hi; i += NT) x[i] = fma(-xp, col[i], x[i]);
                __syncthreads();
                if (tid == 0) x[p] = xp;
            }
        }
        __syncthreads();
        for (int i = tid; i < n; i += NT) xg[(size_t)i * incx] = x[i];
        __syncthreads();
    }
}

// One lane per vector: Left -> column v of B (length m), Right -> row v
// (length n).  The loops are dtrsm.f's, restricted to that vector.
// G lanes per matrix (a power of two; 64 / G matrices share the wavefront when
// a matrix has fewer than 64 vectors).  STAGED: each lane's vector lives in
// LDS as vec[k][lane] for the whole solve (read and written back once).
template <char SIDE, char UPLO, char TRANS, bool STAGED>
__global__ void __launch_bounds__(64) trsm_kernel(int m, int n, double alpha, double **A_array, int lda,
                                                  double **B_array, int ldb, bool unit, int count, int G)
{
    extern __shared__ double vec[];
    const int lane = threadIdx.x;
    const int L = SIDE == 'L' ? m : n;          // vector length
    const int nv = SIDE == 'L' ? n : m;         // vectors per matrix
    const int P = 64 / G;
    const int vblocks = (nv + G - 1) / G;
    const int groups = (count + P - 1) / P;
    for (int item = blockIdx.x; item < groups * vblocks; item += gridDim.x) {
        const int id = (item / vblocks) * P + lane / G;
        const int v = (item % vblocks) * G + lane % G;
        if (id >= count || v >= nv) continue;
        const double *a = A_array[id];
        double *b = B_array[id];
#define AA(i, j) a[(size_t)(i) + (size_t)(j) * lda]
#define GV(k) (SIDE == 'L' ? b[(size_t)(k) + (size_t)v * ldb] : b[(size_t)v + (size_t)(k) * ldb])
#define XV(k) (STAGED ? vec[(size_t)(k) * 64 + lane] : GV(k))
        if (alpha == 0.0) {
            for (int k = 0; k < L; ++k) GV(k) = 0.0;
            continue;
        }
        if (STAGED) {
#pragma unroll 8
            for (int k = 0; k < L; ++k) vec[(size_t)k * 64 + lane] = GV(k);
        }
        if (SIDE == 'L') {
            if (TRANS == 'N') {
                if (alpha != 1.0)
                    #pragma unroll 4
                    for (int i = 0; i < m; ++i) XV(i) = alpha * XV(i);
                if (UPLO == 'U') {
                    for (int k = m - 1; k >= 0; --k) {
                        double bk = XV(k);
                        if (bk != 0.0) {
                            if (!unit) bk = bk / AA(k, k);
                            XV(k) = bk;
                            #pragma unroll 4
                            for (int i = 0; i < k; ++i) XV(i) = fma(-bk, AA(i, k), XV(i));
                        }
                    }
                } else {
                    for (int k = 0; k < m; ++k) {
                        double bk = XV(k);
                        if (bk != 0.0) {
                            if (!unit) bk = bk / AA(k, k);
                            XV(k) = bk;
                            #pragma unroll 4
                            for (int i = k + 1; i < m; ++i) XV(i) = fma(-bk, AA(i, k), XV(i));
                        }
                    }
                }
            } else if (UPLO == 'U') {
                for (int i = 0; i < m; ++i) {
                    double t = alpha * XV(i);
                    #pragma unroll 4
                    for (int k = 0; k < i; ++k) t = fma(-AA(k, i), XV(k), t);
                    if (!unit) t = t / AA(i, i);
                    XV(i) = t;
                }
            } else {
                for (int i = m - 1; i >= 0; --i) {
                    double t = alpha * XV(i);
                    #pragma unroll 4
                    for (int k = i + 1; k < m; ++k) t = fma(-AA(k, i), XV(k), t);
                    if (!unit) t = t / AA(i, i);
                    XV(i) = t;
                }
            }
        } else if (TRANS == 'N') {
            if (UPLO == 'U') {
                for (int j = 0; j < n; ++j) {
                    double t = XV(j);
                    if (alpha != 1.0) t = alpha * t;
                    for (int k = 0; k < j; ++k) {
                        const double akj = AA(k, j);
                        if (akj != 0.0) t = fma(-akj, XV(k), t);
                    }
                    if (!unit) t = (1.0 / AA(j, j)) * t;
                    XV(j) = t;
                }
            } else {
                for (int j = n - 1; j >= 0; --j) {
                    double t = XV(j);
                    if (alpha != 1.0) t = alpha * t;
                    for (int k = j + 1; k < n; ++k) {
                        const double akj = AA(k, j);
                        if (akj != 0.0) t = fma(-akj, XV(k), t);
                    }
                    if (!unit) t = (1.0 / AA(j, j)) * t;
                    XV(j) = t;
                }
            }
        } else if (UPLO == 'U') {
            for (int k = n - 1; k >= 0; --k) {
                double bk = XV(k);
                if (!unit) bk = (1.0 / AA(k, k)) * bk;
                for (int j = 0; j < k; ++j) {
                    const double ajk = AA(j, k);
                    if (ajk != 0.0) XV(j) = fma(-ajk, bk, XV(j));
                }
                if (alpha != 1.0) bk = alpha * bk;
                XV(k) = bk;
            }
        } else {
            for (int k = 0; k < n; ++k) {
                double bk = XV(k);
                if (!unit) bk = (1.0 / AA(k, k)) * bk;
                for (int j = k + 1; j < n; ++j) {
                    const double ajk = AA(j, k);
                    if (ajk != 0.0) XV(j) = fma(-ajk, bk, XV(j));
                }
                if (alpha != 1.0) bk = alpha * bk;
                XV(k) = bk;
            }
        }
        if (STAGED)
            for (int k = 0; k < L; ++k) GV(k) = vec[(size_t)k * 64 + lane];
#undef AA
#undef GV
#undef XV
    }
}

// y := alpha A x + beta y, one thread per row.
__global__ void __launch_bounds__(256) gemv_n_kernel(int m, int n, double alpha, double **A_array, int lda,
                                                    double **x_array, int incx, double beta, double **y_array,
                                                    int incy, int count)
{
    const int rblocks = (m + 255) / 256;
    for (int item = blockIdx.x; item < count * rblocks; item += gridDim.x) {
        const int id = item / rblocks;
        const int i = (item % rblocks) * 256 + (int)threadIdx.x;
        if (i >= m) continue;
        const double *a = A_array[id] + i;
        const double *x = x_array[id];
        double *yp = y_array[id] + (size_t)i * incy;
        double y = (beta == 0.0) ? 0.0 : (beta == 1.0 ? *yp : beta * *yp);
        if (alpha != 0.0) {
            for (int j = 0; j < n; ++j) y = fma(alpha * x[(size_t)j * incx], a[(size_t)j * lda], y);
        }
        *yp = y;
    }
}

// y := alpha A' x + beta y, one thread per column, 64 columns per workgroup,
// A staged 32 rows at a time.
__global__ void __launch_bounds__(64) gemv_t_kernel(int m, int n, double alpha, double **A_array, int lda,
                                                   double **x_array, int incx, double beta, double **y_array,
                                                   int incy, int count)
{
    __shared__ double tile[64][33];
    const int lane = threadIdx.x;
    const int cblocks = (n + 63) / 64;
    for (int item = blockIdx.x; item < count * cblocks; item += gridDim.x) {
        const int id = item / cblocks;
        const int j0 = (item % cblocks) * 64;
        const double *a = A_array[id];
        const double *x = x_array[id];
        const int j = j0 + lane;
        double t = 0.0;
        if (alpha != 0.0) {
            for (int i0 = 0; i0 < m; i0 += 32) {
                const int rows = min(32, m - i0);
                // 64 columns x 32 rows; lanes 0-31 read one column, 32-63 the next
#pragma unroll 4
                for (int k = 0; k < 32; ++k) {
                    const int c = 2 * k + (lane >> 5), r = lane & 31;
                    if (r < rows && j0 + c < n) tile[c][r] = a[(size_t)(i0 + r) + (size_t)(j0 + c) * lda];
                }
                __syncthreads();
                if (j < n)
                    for (int r = 0; r < rows; ++r) t = fma(tile[lane][r], x[(size_t)(i0 + r) * incx], t);
                __syncthreads();
            }
        }
        if (j < n) {
            double *yp = y_array[id] + (size_t)j * incy;
            double y = (beta == 0.0) ? 0.0 : (beta == 1.0 ? *yp : beta * *yp);
            if (alpha != 0.0) y = fma(alpha, t, y);
            *yp = y;
        }
    }
}

// LU without pivoting, one workgroup per matrix: for each column the
// multipliers (reciprocal when |pivot| >= DBL_MIN, else division; a zero
// pivot leaves the column unscaled and sets info), then the rank-1 update
// column by column with dger's skip of zero pivot-row entries, so every
// entry receives its updates as single fmas in increasing k.  LDSM: the
// matrix is held in LDS (m n doubles); otherwise it is updated in place.
template <bool LDSM, int NT>
__global__ void __launch_bounds__(NT) getrf_nopiv_kernel(int m, int n, double **A_array, int lda, int *info,
                                                         int count)
{
    extern __shared__ double sa[];
    const int tid = threadIdx.x;
    const int mn = min(m, n);
    for (int id = blockIdx.x; id < count; id += gridDim.x) {
        double *ga = A_array[id];
        double *a = LDSM ? sa : ga;
        const int ld = LDSM ? m : lda;
        if (LDSM) {
#pragma unroll 8
            for (int c = tid >> 6; c < n; c += NT / 64)
                for (int i = tid & 63; i < m; i += 64) sa[i + c * m] = ga[i + (size_t)c * lda];
            __syncthreads();
        }
        int inf = 0;
#pragma unroll 1
        for (int j = 0; j < mn; ++j) {
            const double p = a[j + (size_t)j * ld];
            if (p != 0.0) {
                const bool rcp = fabs(p) >= DBL_MIN;
                const double r = 1.0 / p;
                for (int i = j + 1 + tid; i < m; i += NT) {
                    double *x = &a[i + (size_t)j * ld];
                    *x = rcp ? r * *x : *x / p;
                }
            } else if (inf == 0) {
                inf = j + 1;
            }
            __syncthreads();
            if (j < mn - 1) {
                // lanes along the rows, the four wavefronts along the columns
#pragma unroll 4
                for (int c = j + 1 + (tid >> 6); c < n; c += NT / 64) {
                    const double y = a[j + (size_t)c * ld];
                    if (y != 0.0)
                        for (int i = j + 1 + (tid & 63); i < m; i += 64)
                            a[i + (size_t)c * ld] = fma(a[i + (size_t)j * ld], -y, a[i + (size_t)c * ld]);
                }
            }
            __syncthreads();
        }
        if (LDSM)
            for (int c = tid >> 6; c < n; c += NT / 64)
                for (int i = tid & 63; i < m; i += 64) ga[i + (size_t)c * lda] = sa[i + c * m];
        if (tid == 0 && info) info[id] = inf;
        __syncthreads();
    }
}

// Random butterfly transformation, depth 2 (magmablas/zgerbt_func_batched.cu):
// A <- U' A V with U = U1 U2, V = V1 V2.  Level 1 acts on the two halves
// (weights n.. of u and v), level 2 on the whole order (weights 0..).  A
// butterfly of order N pairs index x with x + N/2: (r_x, r_y) <- (w_x (r_x + r_y),
// w_y (r_x - r_y)); an odd leftover index is unchanged.  Each level is a pass
// over the rows (every column) and one over the columns (every row), so the
// transform is exact for every n (the reference's kernels update only the
// paired square, consistent only when both orders are multiples of 4).  One
// workgroup per matrix, in place; the passes are separated by barriers.
__global__ void __launch_bounds__(256) rbt_matrix_kernel(int n, double **A_array, int lda, const double *u,
                                                        const double *v, int count)
{
    const int tid = threadIdx.x;
    const int h = n / 2, hh = h / 2;
    for (int id = blockIdx.x; id < count; id += gridDim.x) {
        double *a = A_array[id];
#define AE(i, j) a[(size_t)(i) + (size_t)(j) * lda]
        for (int pass = 0; pass < 4; ++pass) {
            const bool rows = (pass & 1) == 0, lvl1 = pass < 2;
            const int pairs = lvl1 ? hh : h;          // pairs per block
            const int blocks = lvl1 ? 2 : 1;
            const int N2 = lvl1 ? hh : h;             // pair distance
            const int lane = tid & 63, wv = tid >> 6;
            for (int b = 0; b < blocks; ++b) {
                const int o = lvl1 ? b * h : 0;
                const double *w = lvl1 ? ((rows ? u : v) + n + b * h) : (rows ? u : v);
                // lanes run along the contiguous index: the pair index for a
                // rows pass (down a column), the row for a columns pass
                if (rows) {
                    for (int c = wv; c < n; c += 4)
                        for (int x = lane; x < pairs; x += 64) {
                            double *p0 = &AE(o + x, c), *p1 = &AE(o + x + N2, c);
                            const double s0 = *p0, s1 = *p1;
                            *p0 = w[x] * (s0 + s1);
                            *p1 = w[x + N2] * (s0 - s1);
                        }
                } else {
                    for (int x = wv; x < pairs; x += 4)
                        for (int r = lane; r < n; r += 64) {
                            double *p0 = &AE(r, o + x), *p1 = &AE(r, o + x + N2);
                            const double s0 = *p0, s1 = *p1;
                            *p0 = w[x] * (s0 + s1);
                            *p1 = w[x + N2] * (s0 - s1);
                        }
                }
            }
            __syncthreads();
        }
#undef AE
    }
}

// The right-hand sides: forward (b <- U' b: level 1 on the halves, then level
// 2), or backward (x <- V y: level 2, then level 1), every column.
__global__ void __launch_bounds__(256) rbt_vector_kernel(int n, int nrhs, double **B_array, int ldb,
                                                        const double *w, bool forward, int count)
{
    const int tid = threadIdx.x;
    const int h = n / 2, hh = h / 2;
    for (int id = blockIdx.x; id < count; id += gridDim.x) {
        double *bm = B_array[id];
        for (int step = 0; step < 2; ++step) {
            const bool lvl1 = forward ? step == 0 : step == 1;
            const int pairs = lvl1 ? hh : h;
            const int blocks = lvl1 ? 2 : 1;
            const int N2 = lvl1 ? hh : h;
            const int lane = tid & 63, wv = tid >> 6;
            for (int blk = 0; blk < blocks; ++blk) {
                const int o = lvl1 ? blk * h : 0;
                const double *ww = lvl1 ? w + n + blk * h : w;
                for (int j = wv; j < nrhs; j += 4) {
                    double *b = bm + (size_t)j * ldb + o;
                    for (int x = lane; x < pairs; x += 64) {
                        if (forward) {
                            const double a1 = b[x] + b[x + N2], a2 = b[x] - b[x + N2];
                            b[x] = ww[x] * a1;
                            b[x + N2] = ww[x + N2] * a2;
                        } else {
                            const double a1 = ww[x] * b[x], a2 = ww[x + N2] * b[x + N2];
                            b[x] = a1 + a2;
                            b[x + N2] = a1 - a2;
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
}

// The butterfly weights: 2n each of u and v, exp((r - 1/2) / 10) with r in
// [0,1) (the distribution of src/zgerbt_gpu.cpp:16-28) from splitmix64 of
// 2^32 + 2 i (u) and 2^32 + 2 i + 1 (v): a fixed function of n.
uint64_t rbt_mix(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void rbt_butterflies(int n, double *u, double *v)
{
    for (int i = 0; i < 2 * n; ++i) {
        const double ru = (double)(rbt_mix((1ull << 32) + 2ull * i) >> 11) * 0x1.0p-53;
        const double rv = (double)(rbt_mix((1ull << 32) + 2ull * i + 1) >> 11) * 0x1.0p-53;
        u[i] = std::exp((ru - 0.5) / 10.0);
        v[i] = std::exp((rv - 0.5) / 10.0);
    }
}

// Householder QR (the reflectors of netlib dgeqr2: dlarfg's beta = -sign(alpha)
// hypot(alpha, ||x||), tau = (beta - alpha) / beta, v = (1; x / (alpha - beta));
// dlarf applied from the left), one workgroup of 256 per matrix, in place.
// ||x|| is max |x_r| times the root of the sum of (x_r / max)^2, so no square
// over- or underflows.  Per column: two reductions, the reflector, w = C'v
// with one wavefront per column of C (lanes along the rows), C -= tau v w'.
template <bool LDSM>
__global__ void __launch_bounds__(256) geqrf_kernel(int m, int n, double **A_array, int lda, double **tau_array,
                                                   int count)
{
    __shared__ double redm[4], reds[4];
    __shared__ double coef[3];
    extern __shared__ double dyn[];
    double *v = dyn;              // m: the reflector (v[i] = 1)
    double *w = dyn + m;          // n: C'v
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int k = min(m, n);
    for (int id = blockIdx.x; id < count; id += gridDim.x) {
        double *ga = A_array[id];
        double *tau = tau_array[id];
        // LDSM: the matrix is held in LDS (after v and w) for the whole factorization
        double *a = LDSM ? dyn + m + n : ga;
        const int ld = LDSM ? m : lda;
        if (LDSM) {
#pragma unroll 8
            for (int c = wave; c < n; c += 4)
                for (int r = lane; r < m; r += 64) a[r + c * m] = ga[r + (size_t)c * lda];
            __syncthreads();
        }
#define AQ(r, c) a[(size_t)(r) + (size_t)(c) * ld]
#pragma unroll 1
        for (int i = 0; i < k; ++i) {
            // max |x| and sum (x / max)^2: wavefront shuffles, then the four
            // wavefronts' values through LDS
            double mx = 0.0;
            for (int r = i + 1 + tid; r < m; r += 256) mx = fmax(mx, fabs(AQ(r, i)));
            for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
            if (lane == 0) redm[wave] = mx;
            __syncthreads();
            const double amax = fmax(fmax(redm[0], redm[1]), fmax(redm[2], redm[3]));
            double sq = 0.0;
            if (amax > 0.0)
                for (int r = i + 1 + tid; r < m; r += 256) {
                    const double t = AQ(r, i) / amax;
                    sq = fma(t, t, sq);
                }
            for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off);
            if (lane == 0) reds[wave] = sq;
            __syncthreads();
            if (tid == 0) {
                const double alpha = AQ(i, i);
                const double xnorm = amax * sqrt((reds[0] + reds[1]) + (reds[2] + reds[3]));
                double t = 0.0, beta = alpha, sc = 0.0;
                if (xnorm != 0.0) {
                    const double xa = fabs(alpha), wmax = fmax(xa, xnorm), wmin = fmin(xa, xnorm);
                    const double hyp = wmax * sqrt(1.0 + (wmin / wmax) * (wmin / wmax));
                    beta = -copysign(hyp, alpha);
                    t = (beta - alpha) / beta;
                    sc = 1.0 / (alpha - beta);
                }
                coef[0] = beta;
                coef[1] = t;
                coef[2] = sc;
                tau[i] = t;
            }
            __syncthreads();
            const double ti = coef[1], sc = coef[2];
            if (ti != 0.0) {
                for (int r = i + 1 + tid; r < m; r += 256) {
                    const double x = AQ(r, i) * sc;
                    AQ(r, i) = x;
                    v[r] = x;
                }
                if (tid == 0) {
                    AQ(i, i) = coef[0];
                    v[i] = 1.0;
                }
                __syncthreads();
#pragma unroll 4
                for (int c = i + 1 + wave; c < n; c += 4) {
                    double p = 0.0;
                    for (int r = i + lane; r < m; r += 64) p = fma(v[r], AQ(r, c), p);
                    for (int off = 32; off > 0; off >>= 1) p += __shfl_xor(p, off);
                    if (lane == 0) w[c] = p;
                }
                __syncthreads();
#pragma unroll 4
                for (int c = i + 1 + wave; c < n; c += 4) {
                    const double t = -ti * w[c];
                    for (int r = i + lane; r < m; r += 64) AQ(r, c) = fma(t, v[r], AQ(r, c));
                }
            }
            __syncthreads();
        }
        if (LDSM) {
            for (int c = wave; c < n; c += 4)
                for (int r = lane; r < m; r += 64) ga[r + (size_t)c * lda] = a[r + c * m];
            __syncthreads();
        }
#undef AQ
    }
}

// invA <- P' I, the identity with getrf's row interchanges applied in order
// (row i is e_perm(i)', perm tracked in LDS), and info = the first k (1-based)
// with U(k,k) exactly zero, else 0.
__global__ void __launch_bounds__(256) perm_identity_kernel(int n, double **A_array, int lda, int **ipiv_array,
                                                           double **I_array, int ldi, int *info, int count)
{
    extern __shared__ int perm[];
    __shared__ int first;
    const int tid = threadIdx.x;
    for (int id = blockIdx.x; id < count; id += gridDim.x) {
        const int *ipiv = ipiv_array[id];
        if (tid == 0) {
            first = INT_MAX;
            for (int i = 0; i < n; ++i) perm[i] = i;
            for (int k = 0; k < n; ++k) {
                const int p = ipiv[k] - 1;
                if (p >= 0 && p < n && p != k) {
                    const int t = perm[k];
                    perm[k] = perm[p];
                    perm[p] = t;
                }
            }
        }
        __syncthreads();
        const double *a = A_array[id];
        for (int i = tid; i < n; i += 256)
            if (a[i + (size_t)i * lda] == 0.0) atomicMin(&first, i + 1);
        double *w = I_array[id];
        for (int j = tid >> 6; j < n; j += 4)
            for (int i = tid & 63; i < n; i += 64) w[i + (size_t)j * ldi] = (j == perm[i]) ? 1.0 : 0.0;
        __syncthreads();
        if (tid == 0 && info) info[id] = (first == INT_MAX) ? 0 : first;
        __syncthreads();
    }
}

template <bool T, bool U>
magma_int_t launch_trsv(int n, double **A, int lda, double **x, int incx, bool unit, int count, hipStream_t st)
{
    int pb = (int)(kTrsvLds / (sizeof(double) * (size_t)n)) - 1;
    pb = std::max(1, std::min(32, pb));
    const size_t lds = sizeof(double) * (size_t)n * (pb + 1);
    // workgroups resident per CU as the 160 KB of LDS allows
    const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(16, (160 * 1024) / lds));
    if (n <= 128)
        trsv_kernel<T, U, 64><<<persistent_grid(count, per_cu), 64, lds, st>>>(n, A, lda, x, incx, unit, pb, count);
    else
        trsv_kernel<T, U, kTrsvThreads><<<persistent_grid(count, per_cu), kTrsvThreads, lds, st>>>(
            n, A, lda, x, incx, unit, pb, count);
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

template <char S, char U, char T>
magma_int_t launch_trsm(int m, int n, double alpha, double **A, int lda, double **B, int ldb, bool unit, int count,
                        hipStream_t st)
{
    const int L = S == 'L' ? m : n, nv = S == 'L' ? n : m;
    // one right-hand side, side Left, alpha 1: dtrsm's loops for NoTrans (both
    // triangles) and for Trans/Upper are dtrsv's, so the workgroup-per-matrix
    // trsv kernel gives the same bits with the pivots' columns staged in LDS
    if (S == 'L' && n == 1 && alpha == 1.0 && (T == 'N' || U == 'U') && m <= MAGMA_AMD_TRSV_MAX_N)
        return launch_trsv<T == 'T', U == 'U'>(m, A, lda, B, 1, unit, count, st);
    int G = 1;
    while (G < 64 && G < nv) G *= 2;
    const long items = (long)((count + 64 / G - 1) / (64 / G)) * ((nv + G - 1) / G);
    if (L <= kTrsmStagedMax) {
        const size_t lds = sizeof(double) * 64 * (size_t)L;
        trsm_kernel<S, U, T, true><<<persistent_grid(items, 8), 64, lds, st>>>(m, n, alpha, A, lda, B, ldb,
                                                                                       unit, count, G);
    } else {
        trsm_kernel<S, U, T, false><<<persistent_grid(items, 8), 64, 0, st>>>(m, n, alpha, A, lda, B, ldb,
                                                                                      unit, count, G);
    }
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

// Enqueue only (no argument checks, no synchronization): the no-pivot
// factorization, and the two triangular solves of getrs without interchanges
// (reference dtrsm order, as dgetrs does).
magma_int_t enqueue_getrf_nopiv(int m, int n, double **A, int lda, int *info, int count, hipStream_t st)
{
    if (m == 0 || n == 0) return zero_info(info, count, st);
    if (m > 2048 || n > 2048) return MAGMA_ERR_NOT_SUPPORTED;
    const size_t lds = sizeof(double) * (size_t)m * n;
    const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, (160 * 1024) / std::max<size_t>(lds, 1)));
    // (one wavefront per matrix at n <= 64 measured 2.4x slower than four)
    if (lds <= 64 * 1024)
        getrf_nopiv_kernel<true, 256><<<persistent_grid(count, per_cu), 256, lds, st>>>(m, n, A, lda, info, count);
    else
        getrf_nopiv_kernel<false, 256><<<persistent_grid(count, 2), 256, 0, st>>>(m, n, A, lda, info, count);
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

magma_int_t enqueue_getrs_nopiv(bool notrans, int n, int nrhs, double **A, int lda, double **B, int ldb, int count,
                                hipStream_t st)
{
    if (n == 0 || nrhs == 0) return MAGMA_SUCCESS;
    magma_int_t r;
    if (notrans) {
        r = launch_trsm<'L', 'L', 'N'>(n, nrhs, 1.0, A, lda, B, ldb, true, count, st);
        if (r == MAGMA_SUCCESS) r = launch_trsm<'L', 'U', 'N'>(n, nrhs, 1.0, A, lda, B, ldb, false, count, st);
    } else {
        r = launch_trsm<'L', 'U', 'T'>(n, nrhs, 1.0, A, lda, B, ldb, false, count, st);
        if (r == MAGMA_SUCCESS) r = launch_trsm<'L', 'L', 'T'>(n, nrhs, 1.0, A, lda, B, ldb, true, count, st);
    }
    return r;
}

}  // namespace
}  // namespace magma_amd

using namespace magma_amd;

extern "C" {


// include/magma_zbatched.h:237-242.  Argument checks and codes of
// magmablas/ztrsv_batched.cu:82-92, which report n as -5 and lda as -8;
// incb must be >= 1 (reported on stderr; the reference does not check it).
void magmablas_dtrsv_batched(magma_uplo_t uplo, magma_trans_t trans, magma_diag_t diag, magma_int_t n,
                             double **A_array, magma_int_t lda, double **b_array, magma_int_t incb,
                             magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t info = 0;
    if (uplo != MagmaUpper && uplo != MagmaLower) info = -1;
    else if (trans != MagmaNoTrans && trans != MagmaTrans && trans != MagmaConjTrans) info = -2;
    else if (diag != MagmaUnit && diag != MagmaNonUnit) info = -3;
    else if (n < 0) info = -5;
    else if (lda < std::max(1, n)) info = -8;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return;
    }
    if (n == 0 || batchCount <= 0) return;
    if (incb < 1) {
        std::fprintf(stderr, "%s: incb = %d is not supported (must be >= 1)\n", __func__, incb);
        return;
    }
    if (n > MAGMA_AMD_TRSV_MAX_N) {
        std::fprintf(stderr, "%s: n = %d exceeds the supported %d\n", __func__, n, MAGMA_AMD_TRSV_MAX_N);
        return;
    }
    hipStream_t st = stream_of(queue);
    const bool t = trans != MagmaNoTrans, u = uplo == MagmaUpper, unit = diag == MagmaUnit;
    magma_int_t r = t ? (u ? launch_trsv<true, true>(n, A_array, lda, b_array, incb, unit, batchCount, st)
                           : launch_trsv<true, false>(n, A_array, lda, b_array, incb, unit, batchCount, st))
                      : (u ? launch_trsv<false, true>(n, A_array, lda, b_array, incb, unit, batchCount, st)
                           : launch_trsv<false, false>(n, A_array, lda, b_array, incb, unit, batchCount, st));
    if (r == MAGMA_SUCCESS) (void)sync_unless_async(queue);
}

// include/magma_zbatched.h:202-208; argument checks of
// magmablas/ztrsm_batched.cpp:811-829.
void magmablas_dtrsm_batched(magma_side_t side, magma_uplo_t uplo, magma_trans_t transA, magma_diag_t diag,
                             magma_int_t m, magma_int_t n, double alpha, double **dA_array, magma_int_t ldda,
                             double **dB_array, magma_int_t lddb, magma_int_t batchCount, magma_queue_t queue)
{
    const magma_int_t nrowA = side == MagmaLeft ? m : n;
    magma_int_t info = 0;
    if (side != MagmaLeft && side != MagmaRight) info = -1;
    else if (uplo != MagmaUpper && uplo != MagmaLower) info = -2;
    else if (transA != MagmaNoTrans && transA != MagmaTrans && transA != MagmaConjTrans) info = -3;
    else if (diag != MagmaUnit && diag != MagmaNonUnit) info = -4;
    else if (m < 0) info = -5;
    else if (n < 0) info = -6;
    else if (ldda < std::max(1, nrowA)) info = -9;
    else if (lddb < std::max(1, m)) info = -11;
    else if (batchCount < 0) info = -12;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return;
    }
    if (m == 0 || n == 0 || batchCount == 0) return;
    hipStream_t st = stream_of(queue);
    const bool unit = diag == MagmaUnit;
    const bool l = side == MagmaLeft, u = uplo == MagmaUpper, t = transA != MagmaNoTrans;
    magma_int_t r;
#define TRSM(S, U, T) launch_trsm<S, U, T>(m, n, alpha, dA_array, ldda, dB_array, lddb, unit, batchCount, st)
    if (l) r = u ? (t ? TRSM('L', 'U', 'T') : TRSM('L', 'U', 'N')) : (t ? TRSM('L', 'L', 'T') : TRSM('L', 'L', 'N'));
    else   r = u ? (t ? TRSM('R', 'U', 'T') : TRSM('R', 'U', 'N')) : (t ? TRSM('R', 'L', 'T') : TRSM('R', 'L', 'N'));
#undef TRSM
    if (r == MAGMA_SUCCESS) (void)sync_unless_async(queue);
}

// include/magma_zbatched.h:705-712; argument checks of
// magmablas/zgemv_batched_core.cu:112-128.
void magmablas_dgemv_batched(magma_trans_t trans, magma_int_t m, magma_int_t n, double alpha,
                             double **dA_array, magma_int_t ldda, double **dx_array, magma_int_t incx,
                             double beta, double **dy_array, magma_int_t incy, magma_int_t batchCount,
                             magma_queue_t queue)
{
    magma_int_t info = 0;
    if (trans != MagmaNoTrans && trans != MagmaTrans && trans != MagmaConjTrans) info = -1;
    else if (m < 0) info = -2;
    else if (n < 0) info = -3;
    else if (ldda < std::max(1, m)) info = -6;
    else if (incx < 1) info = -8;
    else if (incy < 1) info = -11;
    else if (batchCount < 0) info = -12;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return;
    }
    if (m == 0 || n == 0 || batchCount == 0 || (alpha == 0.0 && beta == 1.0)) return;
    hipStream_t st = stream_of(queue);
    if (trans == MagmaNoTrans) {
        const long items = (long)batchCount * ((m + 255) / 256);
        gemv_n_kernel<<<persistent_grid(items, 4), 256, 0, st>>>(m, n, alpha, dA_array, ldda, dx_array, incx, beta,
                                                                 dy_array, incy, batchCount);
    } else {
        const long items = (long)batchCount * ((n + 63) / 64);
        gemv_t_kernel<<<persistent_grid(items, 8), 64, 0, st>>>(m, n, alpha, dA_array, ldda, dx_array, incx, beta,
                                                                dy_array, incy, batchCount);
    }
    if (hipGetLastError() != hipSuccess) {
        std::fprintf(stderr, "%s: kernel launch failed\n", __func__);
        return;
    }
    (void)sync_unless_async(queue);
}

// ---------------------------------------------------------- no pivoting ----
// include/magma_zbatched.h:585-589; argument checks of
// src/zgetrf_nopiv_batched.cpp:88-99 (m -1, n -2, ldda -4), sizes up to 2048.
magma_int_t magma_dgetrf_nopiv_batched(magma_int_t m, magma_int_t n, double **dA_array, magma_int_t ldda,
                                       magma_int_t *info_array, magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t arg = 0;
    if (m < 0) arg = -1;
    else if (n < 0) arg = -2;
    else if (ldda < std::max(1, m)) arg = -4;
    else if (batchCount < 0) arg = -6;
    if (arg != 0) {
        magma_xerbla(__func__, -arg);
        return arg;
    }
    if (batchCount == 0) return MAGMA_SUCCESS;
    const magma_int_t r = enqueue_getrf_nopiv(m, n, dA_array, ldda, info_array, batchCount, stream_of(queue));
    return r == MAGMA_SUCCESS ? sync_unless_async(queue) : r;
}

// include/magma_zbatched.h:593-598: the two triangular solves of getrs
// without the row interchanges (reference dtrsm order, as dgetrs does).
magma_int_t magma_dgetrs_nopiv_batched(magma_trans_t trans, magma_int_t n, magma_int_t nrhs, double **dA_array,
                                       magma_int_t ldda, double **dB_array, magma_int_t lddb,
                                       magma_int_t *info_array, magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t arg = 0;
    if (trans != MagmaNoTrans && trans != MagmaTrans && trans != MagmaConjTrans) arg = -1;
    else if (n < 0) arg = -2;
    else if (nrhs < 0) arg = -3;
    else if (ldda < std::max(1, n)) arg = -5;
    else if (lddb < std::max(1, n)) arg = -7;
    else if (batchCount < 0) arg = -9;
    if (arg != 0) {
        magma_xerbla(__func__, -arg);
        return arg;
    }
    hipStream_t st = stream_of(queue);
    if (batchCount == 0) return MAGMA_SUCCESS;
    zero_info(info_array, batchCount, st);
    const magma_int_t r = enqueue_getrs_nopiv(trans == MagmaNoTrans, n, nrhs, dA_array, ldda, dB_array, lddb,
                                              batchCount, st);
    return r == MAGMA_SUCCESS ? sync_unless_async(queue) : r;
}

// include/magma_zbatched.h:601-606
magma_int_t magma_dgesv_nopiv_batched(magma_int_t n, magma_int_t nrhs, double **dA_array, magma_int_t ldda,
                                      double **dB_array, magma_int_t lddb, magma_int_t *info_array,
                                      magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t arg = 0;
    if (n < 0) arg = -1;
    else if (nrhs < 0) arg = -2;
    else if (ldda < std::max(1, n)) arg = -4;
    else if (lddb < std::max(1, n)) arg = -6;
    else if (batchCount < 0) arg = -8;
    if (arg != 0) {
        magma_xerbla(__func__, -arg);
        return arg;
    }
    if (batchCount == 0) return MAGMA_SUCCESS;
    hipStream_t st = stream_of(queue);
    magma_int_t r = enqueue_getrf_nopiv(n, n, dA_array, ldda, info_array, batchCount, st);
    if (r == MAGMA_SUCCESS) r = enqueue_getrs_nopiv(true, n, nrhs, dA_array, ldda, dB_array, lddb, batchCount, st);
    return r == MAGMA_SUCCESS ? sync_unless_async(queue) : r;
}

// include/magma_zbatched.h:609-614; src/zgesv_rbt_batched.cpp: A <- U' A V,
// B <- U' B (every column), LU without pivoting, the two triangular solves,
// X = V Y in B.  The butterfly weights are staged through the queue's pinned
// buffer, so the call synchronizes once before its kernels.
magma_int_t magma_dgesv_rbt_batched(magma_int_t n, magma_int_t nrhs, double **dA_array, magma_int_t ldda,
                                    double **dB_array, magma_int_t lddb, magma_int_t *dinfo_array,
                                    magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t arg = 0;
    if (n < 0) arg = -1;
    else if (nrhs < 0) arg = -2;
    else if (ldda < std::max(1, n)) arg = -4;
    else if (lddb < std::max(1, n)) arg = -6;
    else if (batchCount < 0) arg = -8;
    if (arg != 0) {
        magma_xerbla(__func__, -arg);
        return arg;
    }
    hipStream_t st = stream_of(queue);
    if (batchCount == 0) return MAGMA_SUCCESS;
    if (n == 0 || nrhs == 0) {
        zero_info(dinfo_array, batchCount, st);
        return sync_unless_async(queue);
    }
    if (n > 2048) return MAGMA_ERR_NOT_SUPPORTED;
    const size_t bytes = sizeof(double) * 4 * (size_t)n;
    double *host = static_cast<double *>(queue_host(queue, bytes));
    double *dev = static_cast<double *>(queue_workspace(queue, bytes));
    if (!host || !dev) return MAGMA_ERR_DEVICE_ALLOC;
    MAGMA_HIP_CHECK(hipStreamSynchronize(st));      // the pinned buffer may still feed an earlier copy
    rbt_butterflies(n, host, host + 2 * n);
    MAGMA_HIP_CHECK(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, st));
    const double *du = dev, *dv = dev + 2 * n;
    const int grid = persistent_grid(batchCount, 4);
    rbt_matrix_kernel<<<grid, 256, 0, st>>>(n, dA_array, ldda, du, dv, batchCount);
    MAGMA_LAUNCH_CHECK();
    rbt_vector_kernel<<<grid, 256, 0, st>>>(n, nrhs, dB_array, lddb, du, true, batchCount);
    MAGMA_LAUNCH_CHECK();
    magma_int_t r = enqueue_getrf_nopiv(n, n, dA_array, ldda, dinfo_array, batchCount, st);
    if (r == MAGMA_SUCCESS) r = enqueue_getrs_nopiv(true, n, nrhs, dA_array, ldda, dB_array, lddb, batchCount, st);
    if (r != MAGMA_SUCCESS) return r;
    rbt_vector_kernel<<<grid, 256, 0, st>>>(n, nrhs, dB_array, lddb, dv, false, batchCount);
    MAGMA_LAUNCH_CHECK();
    return sync_unless_async(queue);
}

// include/magma_zbatched.h:715-721; argument checks of
// src/zgeqrf_batched.cpp:104-115 (m -1, n -2, lda -4); m, n <= 2048.
// info_array is set to 0 (a QR factorization always exists).
magma_int_t magma_dgeqrf_batched(magma_int_t m, magma_int_t n, double **dA_array, magma_int_t lda,
                                 double **dtau_array, magma_int_t *info_array, magma_int_t batchCount,
                                 magma_queue_t queue)
{
    magma_int_t arg = 0;
    if (m < 0) arg = -1;
    else if (n < 0) arg = -2;
    else if (lda < std::max(1, m)) arg = -4;
    else if (batchCount < 0) arg = -7;
    if (arg != 0) {
        magma_xerbla(__func__, -arg);
        return arg;
    }
    hipStream_t st = stream_of(queue);
    if (batchCount == 0) return MAGMA_SUCCESS;
    zero_info(info_array, batchCount, st);
    if (m == 0 || n == 0) return sync_unless_async(queue);
    if (m > 2048 || n > 2048) return MAGMA_ERR_NOT_SUPPORTED;
    const size_t vw = sizeof(double) * (size_t)(m + n);
    const size_t whole = vw + sizeof(double) * (size_t)m * n;
    if (whole <= 64 * 1024) {
        const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, (160 * 1024) / whole));
        geqrf_kernel<true><<<persistent_grid(batchCount, per_cu), 256, whole, st>>>(m, n, dA_array, lda, dtau_array,
                                                                                    batchCount);
    } else {
        geqrf_kernel<false><<<persistent_grid(batchCount, 4), 256, vw, st>>>(m, n, dA_array, lda, dtau_array,
                                                                             batchCount);
    }
    MAGMA_LAUNCH_CHECK();
    return sync_unless_async(queue);
}

// src/zgetri_outofplace_batched.cpp:82-220.  The reference sets invA = I,
// solves with L (unit) and U, then interchanges the columns in reverse pivot
// order; column j of that is the solution of L U x = e_p(j).  Here invA = P'I
// (row i is e_perm(i)', so column j is e_p(j)) and the two triangular solves
// run on it: the same arithmetic per column as getrs applied to the identity,
// which is bitwise what the reference's sequence computes.  info[i] = k > 0
// reports U(k,k) == 0 (LAPACK dgetri's code); the inverse is still formed, as
// the reference does.
magma_int_t magma_dgetri_outofplace_batched(magma_int_t n, double **dA_array, magma_int_t ldda,
                                            magma_int_t **dipiv_array, double **dinvA_array, magma_int_t lddia,
                                            magma_int_t *info_array, magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t arg = 0;
    if (n < 0) arg = -1;
    else if (ldda < std::max(1, n)) arg = -3;
    else if (lddia < std::max(1, n)) arg = -6;
    else if (batchCount < 0) arg = -8;
    if (arg != 0) {
        magma_xerbla(__func__, -arg);
        return arg;
    }
    hipStream_t st = stream_of(queue);
    if (batchCount == 0) return MAGMA_SUCCESS;
    if (n == 0) {
        zero_info(info_array, batchCount, st);
        return sync_unless_async(queue);
    }
    if (n > 2048) return MAGMA_ERR_NOT_SUPPORTED;
    perm_identity_kernel<<<persistent_grid(batchCount, 4), 256, sizeof(int) * (size_t)n, st>>>(
        n, dA_array, ldda, dipiv_array, dinvA_array, lddia, info_array, batchCount);
    MAGMA_LAUNCH_CHECK();
    magma_int_t r = launch_trsm<'L', 'L', 'N'>(n, n, 1.0, dA_array, ldda, dinvA_array, lddia, true, batchCount, st);
    if (r == MAGMA_SUCCESS)
        r = launch_trsm<'L', 'U', 'N'>(n, n, 1.0, dA_array, ldda, dinvA_array, lddia, false, batchCount, st);
    if (r != MAGMA_SUCCESS) return r;
    return sync_unless_async(queue);
}

}  // extern "C"
