// lu_reg.h -- register-resident factorizations for small matrices
// (max(m,n) <= 64): one wavefront holds 64/NMAX matrices, one matrix row per
// lane, the row's NMAX entries in VGPRs.
//
// Replaces, for these sizes, the reference's per-column launch sequence
// (magma_dgetf2_batched: zcomputecolumn_kernel_shared_batched + zswap +
// getf2trsm + gemm, magmablas/zgetf2_kernels.cu:353-795, src/zgetf2_batched.cpp:97-194)
// and its lpout Cholesky (magmablas/zpotf2_kernels.cu:114-214), with one
// launch and no pass through HBM between columns.
//
// Numerics follow netlib LAPACK (dgetf2 / dpotrf2 with reference BLAS loop
// orders): every element gets its updates one fused multiply-add at a time in
// increasing k, multipliers are scaled by the reciprocal of the pivot (division
// when |pivot| < DBL_MIN), the pivot is the first maximum of |a| with NaN taken
// only in the first position (idamax).  The factors are therefore the same
// bits as the CPU restatement that the tests compare against.
//
// Row interchanges are tracked as logical positions (pos) instead of moving
// registers between lanes: lane r keeps physical row r; the row written to
// logical position pos at the end.  A right-hand side row travels with its
// matrix row, so the fused solve needs no permutation step.
#include <cfloat>
#include <climits>

#pragma once
#include "kernels.h"

namespace magma_amd {
namespace lu_reg_detail {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kRhsChunk = 4;

// Intra-wave LDS hand-off: orders an LDS write by some lanes before LDS reads
// by other lanes of the same wavefront.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int pick(const int *arr, int fixed, int id) { return arr ? arr[id] : fixed; }

// ---------------------------------------------------------------------------
// LU, partial pivoting.  Template NMAX: rows/cols held per lane group (8..64).
// Minimum resident waves per SIMD asked of the register allocator.
template <int NMAX> constexpr int reg_waves_per_eu() { return NMAX >= 64 ? 2 : (NMAX >= 32 ? 3 : 4); }

// Columns per unrolled chunk (see the rotation in getrf_reg_kernel).
template <int NMAX> constexpr int reg_chunk() { return NMAX > 32 ? 8 : NMAX; }

// a[k] <- a[(k + CH) mod NMAX]
template <int NMAX, int CH>
__device__ __forceinline__ void rotate_left(double (&a)[NMAX])
{
    double t[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) t[k] = a[k];
#pragma unroll
    for (int k = 0; k + CH < NMAX; ++k) a[k] = a[k + CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) a[NMAX - CH + k] = t[k];
}

template <int NMAX, bool SOLVE>
__global__ void __launch_bounds__(kWave *kWavesPerBlock)
__attribute__((amdgpu_waves_per_eu(reg_waves_per_eu<NMAX>()))) getrf_reg_kernel(MatSet s)
{
    constexpr int MPW = kWave / NMAX;
    __shared__ double rowbuf[kWavesPerBlock][kWave];
    __shared__ double aux[kWavesPerBlock][MPW][kRhsChunk];

    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const int g = lane / NMAX;
    const int r = lane % NMAX;
    double *buf = &rowbuf[w][g * NMAX];
    double *xb = aux[w][g];

    const int item = (blockIdx.x * kWavesPerBlock + w) * MPW + g;
    const bool live = item < s.count;
    const int id = live ? (s.index ? s.index[item] : item) : 0;
    const int m = live ? pick(s.m_arr, s.m, id) : 0;
    const int n = live ? pick(s.n_arr, s.n, id) : 0;
    const int lda = live ? pick(s.lda_arr, s.lda, id) : 1;
    double *A = live ? s.A[id] : nullptr;
    const int mn = min(m, n);

    double a[NMAX];
#pragma unroll
    for (int k = 0; k < NMAX; ++k) a[k] = (r < m && k < n) ? A[r + (size_t)k * lda] : 0.0;

    int pos = r;
    int myipiv = 0;
    int info = 0;

    // Columns are processed CH at a time.  Between chunks the row is rotated
    // left by CH registers, so column j always sits in register j - jbase and
    // every register index stays a compile-time constant; after NMAX/CH
    // rotations the row is back in place.
    constexpr int CH = reg_chunk<NMAX>();
#pragma unroll 1
    for (int jbase = 0; jbase < NMAX; jbase += CH) {
#pragma unroll
        for (int jj = 0; jj < CH; ++jj) {
            const int j = jbase + jj;
            if (j < mn) {
                // pivot search: first maximum of |a(pos>=j, j)| (netlib idamax)
                const double v = a[jj];
                const bool cand = (r < m) && (pos >= j);
                double key = -1.0;
                if (cand) key = (v != v) ? ((pos == j) ? __builtin_huge_val() : -1.0) : fabs(v);
                int kp = cand ? pos : INT_MAX;
#pragma unroll
                for (int off = NMAX / 2; off > 0; off >>= 1) {
                    const double ok = __shfl_xor(key, off, NMAX);
                    const int op = __shfl_xor(kp, off, NMAX);
                    const bool take = (ok > key) || (ok == key && op < kp);
                    key = take ? ok : key;
                    kp = take ? op : kp;
                }
                const int p = kp;
                // the pivot row (its live entries) to the group's LDS slot
                if (cand && pos == p) {
#pragma unroll
                    for (int k = jj; k < NMAX; ++k) buf[k] = a[k];
                }
                wave_lds_sync();
                const double piv = buf[jj];
                if (r == j) myipiv = p + 1;
                pos = (pos == j) ? p : ((pos == p) ? j : pos);
                const bool below = (r < m) && (pos > j);
                double l = a[jj];
                if (piv != 0.0) {
                    if (below) {
                        if (fabs(piv) >= DBL_MIN) {
                            const double rc = 1.0 / piv;
                            l = rc * l;
                        } else {
                            l = l / piv;
                        }
                    }
                } else if (info == 0) {
                    info = j + 1;
                }
                if (below) a[jj] = l;
                if (j < mn - 1) {
                    // rank-1 update (dger).  netlib skips a column whose pivot-row
                    // entry is exactly zero; that only changes the result when a
                    // multiplier is Inf/NaN, so the exact form runs only then.
                    const int kend = n - jbase;     // live registers are k < kend
                    if (__any(below && !__builtin_isfinite(l))) {
#pragma unroll
                        for (int k = jj + 1; k < NMAX; ++k) {
                            if (k < kend) {
                                const double u = buf[k];
                                if (below && u != 0.0) a[k] = fma(l, -u, a[k]);
                            }
                        }
                    } else {
#pragma unroll
                        for (int k = jj + 1; k < NMAX; ++k) {
                            if (k < kend) {
                                const double u = buf[k];
                                if (below) a[k] = fma(l, -u, a[k]);
                            }
                        }
                    }
                }
                wave_lds_sync();
            }
        }
        if constexpr (CH < NMAX) rotate_left<NMAX, CH>(a);
    }

    if (r < m) {
#pragma unroll
        for (int k = 0; k < NMAX; ++k)
            if (k < n) A[pos + (size_t)k * lda] = a[k];
    }
    if (live && r < mn) s.ipiv[id][r] = myipiv;
    if (live && r == 0) s.info[id] = info;

    if constexpr (SOLVE) {
        // dgetrs('N') on the factors still in registers: forward unit-L sweep
        // and backward U sweep, column-oriented as reference dtrsm.  Skipped for a
        // singular factor (LAPACK dgesv semantics).
        if (!live || info != 0) return;
        const int nrhs = pick(s.nrhs_arr, s.nrhs, id);
        const int ldb = pick(s.ldb_arr, s.ldb, id);
        double *Bm = s.B[id];
        for (int c0 = 0; c0 < nrhs; c0 += kRhsChunk) {
            double y[kRhsChunk];
#pragma unroll
            for (int q = 0; q < kRhsChunk; ++q)
                y[q] = (r < n && c0 + q < nrhs) ? Bm[r + (size_t)(c0 + q) * ldb] : 0.0;
#pragma unroll
            for (int k = 0; k < NMAX; ++k) {
                if (k < n) {
                    if (pos == k && r < n) {
#pragma unroll
                        for (int q = 0; q < kRhsChunk; ++q) xb[q] = y[q];
                    }
                    wave_lds_sync();
                    if (r < n && pos > k) {
#pragma unroll
                        for (int q = 0; q < kRhsChunk; ++q) y[q] = fma(-xb[q], a[k], y[q]);
                    }
                    wave_lds_sync();
                }
            }
#pragma unroll
            for (int k = NMAX - 1; k >= 0; --k) {
                if (k < n) {
                    if (pos == k && r < n) {
#pragma unroll
                        for (int q = 0; q < kRhsChunk; ++q) {
                            if (y[q] != 0.0) y[q] = y[q] / a[k];
                            xb[q] = y[q];
                        }
                    }
                    wave_lds_sync();
                    if (r < n && pos < k) {
#pragma unroll
                        for (int q = 0; q < kRhsChunk; ++q) y[q] = fma(-xb[q], a[k], y[q]);
                    }
                    wave_lds_sync();
                }
            }
#pragma unroll
            for (int q = 0; q < kRhsChunk; ++q)
                if (r < n && c0 + q < nrhs) Bm[pos + (size_t)(c0 + q) * ldb] = y[q];
        }
    }
}

// ---------------------------------------------------------------------------
// Cholesky, lower: lane r holds row r of L (entries k <= r).
template <int NMAX, bool SOLVE>
__global__ void __launch_bounds__(kWave *kWavesPerBlock)
__attribute__((amdgpu_waves_per_eu(reg_waves_per_eu<NMAX>()))) potrf_reg_kernel(MatSet s)
{
    constexpr int MPW = kWave / NMAX;
    __shared__ double colbuf[kWavesPerBlock][kWave];
    __shared__ double aux[kWavesPerBlock][MPW][kRhsChunk];

    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const int g = lane / NMAX;
    const int r = lane % NMAX;
    double *buf = &colbuf[w][g * NMAX];
    double *xb = aux[w][g];

    const int item = (blockIdx.x * kWavesPerBlock + w) * MPW + g;
    const bool live = item < s.count;
    const int id = live ? (s.index ? s.index[item] : item) : 0;
    const int n = live ? pick(s.n_arr, s.n, id) : 0;
    const int lda = live ? pick(s.lda_arr, s.lda, id) : 1;
    double *A = live ? s.A[id] : nullptr;

    double a[NMAX];
#pragma unroll
    for (int k = 0; k < NMAX; ++k) a[k] = (r < n && k <= r) ? A[r + (size_t)k * lda] : 0.0;

    int info = 0;
    constexpr int CH = reg_chunk<NMAX>();
#pragma unroll 1
    for (int jbase = 0; jbase < NMAX; jbase += CH) {
#pragma unroll
        for (int jj = 0; jj < CH; ++jj) {
            const int j = jbase + jj;
            if (j < n && info == 0) {
                if (r == j) xb[0] = a[jj];
                wave_lds_sync();
                const double ajj = xb[0];
                if (!(ajj > 0.0)) {
                    info = j + 1;                    // dpotrf2: a_jj <= 0 or NaN
                } else {
                    const double ljj = sqrt(ajj);
                    const double rc = 1.0 / ljj;
                    const bool below = (r > j) && (r < n);
                    double l = a[jj];
                    if (r == j) a[jj] = ljj;
                    if (below) {
                        l = rc * l;
                        a[jj] = l;
                        buf[r] = l;
                    }
                    wave_lds_sync();
                    // trailing update a_rc -= l_rj * l_cj (dsyrk), c in (j, n);
                    // register k holds column jbase + k, whose l_cj is buf[jbase + k]
                    const int kend = n - jbase;
                    if (__any(below && !__builtin_isfinite(l))) {
#pragma unroll
                        for (int k = jj + 1; k < NMAX; ++k) {
                            if (k < kend) {
                                const double u = buf[jbase + k];
                                if (below && u != 0.0) a[k] = fma(-u, l, a[k]);
                            }
                        }
                    } else {
#pragma unroll
                        for (int k = jj + 1; k < NMAX; ++k) {
                            if (k < kend) {
                                const double u = buf[jbase + k];
                                if (below) a[k] = fma(-u, l, a[k]);
                            }
                        }
                    }
                }
                wave_lds_sync();
            }
        }
        if constexpr (CH < NMAX) rotate_left<NMAX, CH>(a);
    }

    if (r < n) {
#pragma unroll
        for (int k = 0; k < NMAX; ++k)
            if (k <= r) A[r + (size_t)k * lda] = a[k];
    }
    if (live && r == 0) s.info[id] = info;

    if constexpr (SOLVE) {
        // dpotrs('L'): L y = b (column sweep, division by l_kk), then L' x = y.
        // The L' sweep runs column-oriented (row k of L handed to the lanes
        // through LDS), so its sums associate in decreasing k where the
        // reference dtrsm('L','L','T') accumulates a dot product in increasing
        // k; the two agree to rounding.
        if (!live || info != 0) return;
        const int nrhs = pick(s.nrhs_arr, s.nrhs, id);
        const int ldb = pick(s.ldb_arr, s.ldb, id);
        double *Bm = s.B[id];
        for (int c0 = 0; c0 < nrhs; c0 += kRhsChunk) {
            double y[kRhsChunk];
#pragma unroll
            for (int q = 0; q < kRhsChunk; ++q)
                y[q] = (r < n && c0 + q < nrhs) ? Bm[r + (size_t)(c0 + q) * ldb] : 0.0;
#pragma unroll
            for (int k = 0; k < NMAX; ++k) {
                if (k < n) {
                    if (r == k) {
#pragma unroll
                        for (int q = 0; q < kRhsChunk; ++q) {
                            if (y[q] != 0.0) y[q] = y[q] / a[k];
                            xb[q] = y[q];
                        }
                    }
                    wave_lds_sync();
                    if (r > k && r < n) {
#pragma unroll
                        for (int q = 0; q < kRhsChunk; ++q) y[q] = fma(-xb[q], a[k], y[q]);
                    }
                    wave_lds_sync();
                }
            }
#pragma unroll
            for (int k = NMAX - 1; k >= 0; --k) {
                if (k < n) {
                    if (r == k) {
#pragma unroll
                        for (int q = 0; q < kRhsChunk; ++q) {
                            y[q] = y[q] / a[k];
                            xb[q] = y[q];
                        }
#pragma unroll
                        for (int c = 0; c < k; ++c) buf[c] = a[c];
                    }
                    wave_lds_sync();
                    if (r < k) {
                        const double lk = buf[r];
#pragma unroll
                        for (int q = 0; q < kRhsChunk; ++q) y[q] = fma(-lk, xb[q], y[q]);
                    }
                    wave_lds_sync();
                }
            }
#pragma unroll
            for (int q = 0; q < kRhsChunk; ++q)
                if (r < n && c0 + q < nrhs) Bm[r + (size_t)(c0 + q) * ldb] = y[q];
        }
    }
}

template <int NMAX, bool SOLVE>
magma_int_t launch_getrf_reg(const MatSet &s, hipStream_t st)
{
    constexpr int per_block = kWavesPerBlock * (kWave / NMAX);
    const int grid = ceil_div(s.count, per_block);
    getrf_reg_kernel<NMAX, SOLVE><<<grid, kWave * kWavesPerBlock, 0, st>>>(s);
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

template <int NMAX, bool SOLVE>
magma_int_t launch_potrf_reg(const MatSet &s, hipStream_t st)
{
    constexpr int per_block = kWavesPerBlock * (kWave / NMAX);
    const int grid = ceil_div(s.count, per_block);
    potrf_reg_kernel<NMAX, SOLVE><<<grid, kWave * kWavesPerBlock, 0, st>>>(s);
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

}  // namespace lu_reg_detail

// One translation unit per NMAX (lu_reg_<NMAX>.hip) instantiates these.
template <int NMAX> magma_int_t launch_getrf_reg_n(const MatSet &s, bool solve, hipStream_t st);
template <int NMAX> magma_int_t launch_potrf_reg_n(const MatSet &s, bool solve, hipStream_t st);

}  // namespace magma_amd

#define MAGMA_AMD_LU_REG_INSTANTIATE(N)                                                         \
    namespace magma_amd {                                                                       \
    template <> magma_int_t launch_getrf_reg_n<N>(const MatSet &s, bool solve, hipStream_t st) \
    {                                                                                           \
        return solve ? lu_reg_detail::launch_getrf_reg<N, true>(s, st)                          \
                     : lu_reg_detail::launch_getrf_reg<N, false>(s, st);                        \
    }                                                                                           \
    template <> magma_int_t launch_potrf_reg_n<N>(const MatSet &s, bool solve, hipStream_t st) \
    {                                                                                           \
        return solve ? lu_reg_detail::launch_potrf_reg<N, true>(s, st)                          \
                     : lu_reg_detail::launch_potrf_reg<N, false>(s, st);                        \
    }                                                                                           \
    }
