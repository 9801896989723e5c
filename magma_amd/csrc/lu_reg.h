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
//
// Control flow is kept wave-uniform inside the column loops: sizes are kernel
// arguments (scalars) for fixed-size batches, entries outside the matrix are
// zero-padded instead of guarded, and the only per-lane condition of an update
// ("row below the pivot") masks the whole row update once.
#pragma once

#include <cfloat>
#include <climits>

#include "kernels.h"
#include "wave_reduce.h"

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

__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Minimum resident waves per SIMD asked of the register allocator.
template <int NMAX> constexpr int reg_waves_per_eu() { return NMAX >= 64 ? 2 : 4; }
// The n = 64 Cholesky of fixed-size batches fits three (168 registers, a few
// spilled outside the column loops): 20,000 posv with 4 right-hand sides
// measured 0.87 -> 0.75 ms against two.  The variable-size form spills
// heavily at three and keeps two.
template <int NMAX, bool VAR> constexpr int potrf_waves_per_eu()
{
    return (NMAX >= 64 && !VAR) ? 3 : reg_waves_per_eu<NMAX>();
}

// Columns per unrolled chunk.  All columns are unrolled: a rotated frame
// (chunks of 8 with the row rotated between chunks, see rotate_left) keeps
// the code 8 times smaller at NMAX = 64 but needs twice the registers there,
// which spilled the row to scratch.
template <int NMAX> constexpr int reg_chunk() { return NMAX; }

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

// p[id] read as a global (address space 1) pointer and converted back: the
// compiler then emits global rather than flat accesses through it.  A flat
// load also counts against the LDS wait counter, so the first LDS hand-off of
// the column loop would wait for every entry of the row, where a global load
// is waited for only when its register is first used.
__device__ __forceinline__ double *global_ptr(double *const *p, int id)
{
    return (double *)((GDoublePtr const *)p)[id];
}

// Per-item sizes: scalars for fixed-size batches (VAR = false), per-matrix
// device arrays for variable-size batches.
template <bool VAR>
struct Item {
    bool live;
    int id, m, n, lda;
    __device__ Item(const MatSet &s, int item, int cnt, const int *idx)
    {
        live = item < cnt;
        if constexpr (VAR) {
            id = live ? (idx ? idx[item] : item) : 0;
            n = live ? s.n_arr[id] : 0;
            m = live ? (s.m_arr ? s.m_arr[id] : n) : 0;
            lda = live ? s.lda_arr[id] : 1;
        } else {
            id = (live && idx) ? idx[item] : item;
            m = s.m;
            n = s.n;
            lda = s.lda;
        }
    }
};

// Update of the live registers k in (jj, NMAX) (and k < kmax, a wave-uniform
// bound used by the rotated frame): a[k] = fma(l, -u_k, a[k]) with u from LDS
// (NEGU_FIRST: fma(-u_k, l, a[k]), the dsyrk operand order).  EXACT keeps
// netlib's skip of a zero pivot-row entry.
// jj is a constant after the caller's column loop is unrolled.
template <int NMAX, bool EXACT, bool NEGU_FIRST>
__device__ __forceinline__ void row_update(double (&a)[NMAX], int jj, double l, const double *u, int kmax)
{
#pragma unroll
    for (int kc = 0; kc < NMAX; kc += 8) {
        if (kc + 7 <= jj) continue;                 // compile-time: nothing live here
        if (kc >= kmax) break;                      // wave-uniform
#pragma unroll
        for (int k = kc; k < kc + 8; ++k) {
            if (k <= jj) continue;
            const double uk = u[k];
            double r;
            if constexpr (NEGU_FIRST) r = fma(-uk, l, a[k]);
            else r = fma(l, -uk, a[k]);
            if constexpr (EXACT) a[k] = (uk != 0.0) ? r : a[k];
            else a[k] = r;
        }
        // NMAX = 64 holds 128 VGPRs of matrix: stop the scheduler from
        // hoisting every chunk's LDS loads ahead of the FMAs
        if constexpr (NMAX > 32) __builtin_amdgcn_sched_barrier(0);
    }
}

// The same update for NMAX = 64 (one matrix per wavefront) with u_k taken from
// register k of the pivot row's lane src by readlane instead of LDS (at n = 64
// the LDS hand-off of the pivot row measured slower: 0.78 against 0.99 ms for
// 20,000 factorizations).  jj is a constant after the caller's column loop is
// unrolled.
template <int NMAX, bool EXACT>
__device__ __forceinline__ void row_update_from_lane(double (&a)[NMAX], int jj, double l, int src)
{
#pragma unroll
    for (int k = 0; k < NMAX; ++k) {
        if (k <= jj) continue;
        const double uk = readlane_f64(a[k], src);
        const double r = fma(l, -uk, a[k]);
        if constexpr (EXACT) a[k] = (uk != 0.0) ? r : a[k];
        else a[k] = r;
    }
}

// dgetrs('N') on the factors still in registers for right-hand sides
// [c0, c0 + RC) of B (those < nrhs): forward unit-L sweep and backward U
// sweep, column-oriented as reference dtrsm.
template <int NMAX, int RC>
__device__ __forceinline__ void getrs_reg(const double (&a)[NMAX], int pos, int r, int n, double *Bm, int ldb, int c0,
                                          int nrhs, double *xb)
{
    const bool rown = r < n;
    double y[RC];
#pragma unroll
    for (int q = 0; q < RC; ++q)
        y[q] = (rown && c0 + q < nrhs) ? Bm[r + (size_t)(c0 + q) * ldb] : 0.0;
#pragma unroll
    for (int k = 0; k < NMAX; ++k) {
        if (k < n) {
            double yk[RC];
            if constexpr (NMAX == 64) {
                const unsigned long long ow = __builtin_amdgcn_ballot_w64(pos == k && rown);
                const int ol = ow ? (int)__builtin_ctzll(ow) : 0;
#pragma unroll
                for (int q = 0; q < RC; ++q) yk[q] = readlane_f64(y[q], ol);
            } else {
                if (pos == k && rown) {
#pragma unroll
                    for (int q = 0; q < RC; ++q) xb[q] = y[q];
                }
                wave_lds_sync();
#pragma unroll
                for (int q = 0; q < RC; ++q) yk[q] = xb[q];
            }
            if (rown && pos > k) {
#pragma unroll
                for (int q = 0; q < RC; ++q) y[q] = fma(-yk[q], a[k], y[q]);
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
#pragma unroll
    for (int k = NMAX - 1; k >= 0; --k) {
        if (k < n) {
            double yk[RC];
            if constexpr (NMAX == 64) {
                const unsigned long long ow = __builtin_amdgcn_ballot_w64(pos == k && rown);
                const int ol = ow ? (int)__builtin_ctzll(ow) : 0;
                if (pos == k && rown) {
#pragma unroll
                    for (int q = 0; q < RC; ++q)
                        if (y[q] != 0.0) y[q] = y[q] / a[k];
                }
#pragma unroll
                for (int q = 0; q < RC; ++q) yk[q] = readlane_f64(y[q], ol);
            } else {
                if (pos == k && rown) {
#pragma unroll
                    for (int q = 0; q < RC; ++q) {
                        if (y[q] != 0.0) y[q] = y[q] / a[k];
                        xb[q] = y[q];
                    }
                }
                wave_lds_sync();
#pragma unroll
                for (int q = 0; q < RC; ++q) yk[q] = xb[q];
            }
            if (rown && pos < k) {
#pragma unroll
                for (int q = 0; q < RC; ++q) y[q] = fma(-yk[q], a[k], y[q]);
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
#pragma unroll
    for (int q = 0; q < RC; ++q)
        if (rown && c0 + q < nrhs) Bm[pos + (size_t)(c0 + q) * ldb] = y[q];
}

// ---------------------------------------------------------------------------
// LU, partial pivoting.
template <int NMAX, bool SOLVE, bool VAR>
__global__ void __launch_bounds__(kWave *kWavesPerBlock)
__attribute__((amdgpu_waves_per_eu(reg_waves_per_eu<NMAX>()))) getrf_reg_kernel(MatSet s)
{
    constexpr int MPW = kWave / NMAX;
    constexpr int CH = reg_chunk<NMAX>();
    __shared__ double rowbuf[kWavesPerBlock][kWave];
    __shared__ double aux[kWavesPerBlock][MPW][kRhsChunk];

    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const int g = lane / NMAX;
    const int r = lane % NMAX;
    double *buf = &rowbuf[w][g * NMAX];
    double *xb = aux[w][g];

    // persistent over blocks of work: the grid may be smaller than the work
    // (device-resident bucket counts)
    const int cnt = item_count(s);
    const int *idx = item_index(s);
    for (int blk = blockIdx.x; blk * (kWavesPerBlock * MPW) < cnt; blk += gridDim.x) {
        const Item<VAR> it(s, (blk * kWavesPerBlock + w) * MPW + g, cnt, idx);
        const int m = it.m, n = it.n, lda = it.lda;
        const bool rowok = it.live && r < m;
        double *A = it.live ? global_ptr(s.A, it.id) : nullptr;
        const int mn = min(m, n);

        double a[NMAX];
    #pragma unroll
        for (int k = 0; k < NMAX; ++k) a[k] = 0.0;
        if (rowok) {
    #pragma unroll
            for (int k = 0; k < NMAX; ++k)
                if (k < n) a[k] = A[r + (size_t)k * lda];
        }

        int pos = r;
        int myipiv = 0;
        int info = 0;

    #pragma unroll 1
        for (int jbase = 0; jbase < NMAX; jbase += CH) {
            const int kmax = NMAX - jbase;          // registers >= kmax hold finished columns
    #pragma unroll
            for (int jj = 0; jj < CH; ++jj) {
                const int j = jbase + jj;
                if (j < mn) {
                    // pivot search: first maximum of |a(pos>=j, j)| (netlib idamax)
                    const double v = a[jj];
                    const bool cand = rowok && pos >= j;
                    // ordered integer key (0: not a candidate, or NaN past the
                    // first position, which idamax never takes)
                    unsigned long long key = 0;
                    if (cand && (v == v || pos == j)) key = pivot_key(v == v ? fabs(v) : __builtin_huge_val());
                    int kp = cand ? pos : INT_MAX;
                    group_argmax_key<NMAX>(key, kp, lane);
                    const int p = kp;
                    // the pivot row: by readlane from its lane (n = 64), else
                    // through LDS (its lane writes, all read)
                    double piv;
                    int ol = 0;
                    if constexpr (NMAX == 64) {
                        const unsigned long long ow = __builtin_amdgcn_ballot_w64(cand && pos == p);
                        ol = ow ? (int)__builtin_ctzll(ow) : 0;
                        piv = readlane_f64(a[jj], ol);
                    } else {
                        if (cand && pos == p) {
    #pragma unroll
                            for (int k = jj; k < NMAX; ++k)
                                if (k < kmax) buf[k] = a[k];
                        }
                        wave_lds_sync();
                        piv = buf[jj];
                    }
                    if (r == j) myipiv = p + 1;
                    pos = (pos == j) ? p : ((pos == p) ? j : pos);
                    const bool below = rowok && pos > j;
                    double l = a[jj];
                    if (piv == 0.0) {
                        if (info == 0) info = j + 1;
                    } else if (fabs(piv) >= DBL_MIN) {
                        const double rc = 1.0 / piv;
                        l = rc * l;
                    } else {
                        l = l / piv;
                    }
                    if (below) {
                        a[jj] = l;
                        // rank-1 update (dger).  netlib skips a column whose pivot-row
                        // entry is exactly zero; that changes the result only when a
                        // multiplier is Inf/NaN, so the exact form runs only then.
                        if constexpr (NMAX == 64) {
                            if (!__builtin_isfinite(l)) row_update_from_lane<NMAX, true>(a, jj, l, ol);
                            else row_update_from_lane<NMAX, false>(a, jj, l, ol);
                        } else {
                            if (!__builtin_isfinite(l)) row_update<NMAX, true, false>(a, jj, l, buf, kmax);
                            else row_update<NMAX, false, false>(a, jj, l, buf, kmax);
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
            if constexpr (CH < NMAX) rotate_left<NMAX, CH>(a);
        }

        if (rowok) {
    #pragma unroll
            for (int k = 0; k < NMAX; ++k)
                if (k < n) A[pos + (size_t)k * lda] = a[k];
        }
        if (it.live && r < mn) s.ipiv[it.id][r] = myipiv;
        if (it.live && r == 0) s.info[it.id] = info;

        if constexpr (SOLVE) {
            // dgetrs('N') on the factors still in registers: forward unit-L sweep
            // and backward U sweep, column-oriented as reference dtrsm.  Skipped for
            // a singular factor (LAPACK dgesv semantics).
            if (!it.live || info != 0) continue;
            const int nrhs = s.nrhs_arr ? s.nrhs_arr[it.id] : s.nrhs;
            const int ldb = s.ldb_arr ? s.ldb_arr[it.id] : s.ldb;
            double *Bm = global_ptr(s.B, it.id);
            // one right-hand side at a time when the system has one (a group
            // of four otherwise would carry three columns of zeros)
            if (nrhs == 1) getrs_reg<NMAX, 1>(a, pos, r, n, Bm, ldb, 0, 1, xb);
            else
                for (int c0 = 0; c0 < nrhs; c0 += kRhsChunk) getrs_reg<NMAX, kRhsChunk>(a, pos, r, n, Bm, ldb, c0, nrhs, xb);
        }
    }
}

// ---------------------------------------------------------------------------
// Cholesky, lower: lane r holds row r of L (entries k <= r; entries k > r are
// zero-padded scratch and never stored).
template <int NMAX, bool SOLVE, bool VAR>
__global__ void __launch_bounds__(kWave *kWavesPerBlock)
__attribute__((amdgpu_waves_per_eu(potrf_waves_per_eu<NMAX, VAR>()))) potrf_reg_kernel(MatSet s)
{
    constexpr int MPW = kWave / NMAX;
    constexpr int CH = reg_chunk<NMAX>();
    __shared__ double colbuf[kWavesPerBlock][kWave];
    __shared__ double aux[kWavesPerBlock][MPW][kRhsChunk];
    // NMAX = 64: rows of L handed to the L' sweep, kLRows at a time
    constexpr int kLRows = 8;
    __shared__ double lrows[kWavesPerBlock][NMAX == 64 ? kLRows : 1][kWave];

    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const int g = lane / NMAX;
    const int r = lane % NMAX;
    double *buf = &colbuf[w][g * NMAX];
    double *xb = aux[w][g];

    // persistent over blocks of work: the grid may be smaller than the work
    // (device-resident bucket counts)
    const int cnt = item_count(s);
    const int *idx = item_index(s);
    for (int blk = blockIdx.x; blk * (kWavesPerBlock * MPW) < cnt; blk += gridDim.x) {
        const Item<VAR> it(s, (blk * kWavesPerBlock + w) * MPW + g, cnt, idx);
        const int n = it.n, lda = it.lda;
        const bool rowok = it.live && r < n;
        double *A = it.live ? global_ptr(s.A, it.id) : nullptr;

        double a[NMAX];
    #pragma unroll
        for (int k = 0; k < NMAX; ++k) a[k] = 0.0;
        if (rowok) {
    #pragma unroll
            for (int k = 0; k < NMAX; ++k)
                if (k < n && k <= r) a[k] = A[r + (size_t)k * lda];
        }
        // LDS slots of rows that never publish a column must hold finite values:
        // the padded update reads them
        buf[r] = 0.0;

        int info = 0;
        double rdg = 1.0;                        // 1 / l_rr of this lane's row
    #pragma unroll 1
        for (int jbase = 0; jbase < NMAX; jbase += CH) {
            const int kmax = NMAX - jbase;
    #pragma unroll
            for (int jj = 0; jj < CH; ++jj) {
                const int j = jbase + jj;
                if (j < n && info == 0) {
                    double ajj;
                    if constexpr (NMAX == 64) {
                        ajj = readlane_f64(a[jj], j);
                    } else {
                        if (r == j) xb[0] = a[jj];
                        wave_lds_sync();
                        ajj = xb[0];
                    }
                    if (!(ajj > 0.0)) {
                        info = j + 1;                    // dpotrf2: a_jj <= 0 or NaN
                    } else {
                        const double ljj = sqrt(ajj);
                        const double rc = 1.0 / ljj;
                        const bool below = rowok && r > j;
                        const double l = rc * a[jj];
                        if (r == j) {
                            a[jj] = ljj;
                            rdg = rc;
                        }
                        // the multipliers reach the other lanes through LDS (one
                        // write per lane, broadcast reads); at n = 64 that issues
                        // a third of the VALU work of readlane pairs and measured
                        // 1.17 -> 1.09 ms on 20,000 n = 64 posv
                        {
                            if (below) {
                                a[jj] = l;
                                buf[r] = l;
                            }
                            wave_lds_sync();
                            // trailing update a_rc -= l_rj * l_cj (dsyrk), c in (j, n);
                            // register k holds column jbase + k, whose l_cj is buf[jbase + k]
                            if (below) {
                                if (!__builtin_isfinite(l)) row_update<NMAX, true, true>(a, jj, l, buf + jbase, kmax);
                                else row_update<NMAX, false, true>(a, jj, l, buf + jbase, kmax);
                            }
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
            if constexpr (CH < NMAX) rotate_left<NMAX, CH>(a);
        }

        if (rowok) {
    #pragma unroll
            for (int k = 0; k < NMAX; ++k)
                if (k <= r) A[r + (size_t)k * lda] = a[k];
        }
        if (it.live && r == 0) s.info[it.id] = info;

        if constexpr (SOLVE) {
            // dpotrs('L'): L y = b (column sweep), then L' x = y.  Both sweeps
            // scale by the reciprocal 1 / l_kk that the factorization formed
            // for the multipliers (one multiply per right-hand side where the
            // reference dtrsm divides: a division is a dozen dependent VALU
            // instructions, issued by the whole wavefront for one lane, and
            // was most of the solve's issue time).  The L' sweep runs
            // column-oriented (row k of L handed to the lanes through LDS), so
            // its sums associate in decreasing k where the reference
            // dtrsm('L','L','T') accumulates a dot product in increasing k.
            // Both agree with the reference to rounding.
            if (!it.live || info != 0) continue;
            const int nrhs = s.nrhs_arr ? s.nrhs_arr[it.id] : s.nrhs;
            const int ldb = s.ldb_arr ? s.ldb_arr[it.id] : s.ldb;
            double *Bm = global_ptr(s.B, it.id);
            for (int c0 = 0; c0 < nrhs; c0 += kRhsChunk) {
                double y[kRhsChunk];
    #pragma unroll
                for (int q = 0; q < kRhsChunk; ++q)
                    y[q] = (rowok && c0 + q < nrhs) ? Bm[r + (size_t)(c0 + q) * ldb] : 0.0;
    #pragma unroll
                for (int k = 0; k < NMAX; ++k) {
                    if (k < n) {
                        double yk[kRhsChunk];
                        if (r == k) {
    #pragma unroll
                            for (int q = 0; q < kRhsChunk; ++q) {
                                y[q] = y[q] * rdg;
                                xb[q] = y[q];
                            }
                        }
                        // y_k to the other lanes through LDS (at n = 64 as well:
                        // readlane pairs measured 3-4% slower on 20,000 posv)
                        wave_lds_sync();
    #pragma unroll
                        for (int q = 0; q < kRhsChunk; ++q) yk[q] = xb[q];
                        if (rowok && r > k) {
    #pragma unroll
                            for (int q = 0; q < kRhsChunk; ++q) y[q] = fma(-yk[q], a[k], y[q]);
                        }
                        __builtin_amdgcn_wave_barrier();
                    }
                }
                if constexpr (NMAX == 64) {
                    // the same sweep, rolled: rows of L go to LDS kLRows at a
                    // time (every lane of the group writes its own row at once)
                    // and x_k reaches the other lanes through LDS
    #pragma unroll 1
                    for (int kb = ((n - 1) / kLRows) * kLRows; kb >= 0; kb -= kLRows) {
                        if (rowok && r >= kb && r < kb + kLRows) {
    #pragma unroll
                            for (int c = 0; c < NMAX; ++c)
                                if (c < r) lrows[w][r - kb][c] = a[c];
                        }
                        wave_lds_sync();
    #pragma unroll 1
                        for (int k = min(kb + kLRows - 1, n - 1); k >= kb; --k) {
                            if (r == k) {
    #pragma unroll
                                for (int q = 0; q < kRhsChunk; ++q) {
                                    y[q] = y[q] * rdg;
                                    xb[q] = y[q];
                                }
                            }
                            wave_lds_sync();
                            double yk[kRhsChunk];
    #pragma unroll
                            for (int q = 0; q < kRhsChunk; ++q) yk[q] = xb[q];
                            if (rowok && r < k) {
                                const double lk = lrows[w][k - kb][r];
    #pragma unroll
                                for (int q = 0; q < kRhsChunk; ++q) y[q] = fma(-lk, yk[q], y[q]);
                            }
                        }
                        wave_lds_sync();
                    }
                } else {
    #pragma unroll
                for (int k = NMAX - 1; k >= 0; --k) {
                    if (k < n) {
                        if (r == k) {
    #pragma unroll
                            for (int q = 0; q < kRhsChunk; ++q) {
                                y[q] = y[q] * rdg;
                                xb[q] = y[q];
                            }
    #pragma unroll
                            for (int c = 0; c < k; ++c) buf[c] = a[c];
                        }
                        wave_lds_sync();
                        if (rowok && r < k) {
                            const double lk = buf[r];
    #pragma unroll
                            for (int q = 0; q < kRhsChunk; ++q) y[q] = fma(-lk, xb[q], y[q]);
                        }
                        __builtin_amdgcn_wave_barrier();
                    }
                }
                }
    #pragma unroll
                for (int q = 0; q < kRhsChunk; ++q)
                    if (rowok && c0 + q < nrhs) Bm[r + (size_t)(c0 + q) * ldb] = y[q];
            }
        }
    }
}

template <int NMAX, bool SOLVE, bool VAR>
magma_int_t launch_getrf_reg(const MatSet &s, hipStream_t st)
{
    constexpr int per_block = kWavesPerBlock * (kWave / NMAX);
    const int blocks = ceil_div(s.count, per_block);
    const int grid = s.dev_count ? persistent_grid(blocks, 8) : blocks;
    getrf_reg_kernel<NMAX, SOLVE, VAR><<<grid, kWave * kWavesPerBlock, 0, st>>>(s);
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

template <int NMAX, bool SOLVE, bool VAR>
magma_int_t launch_potrf_reg(const MatSet &s, hipStream_t st)
{
    constexpr int per_block = kWavesPerBlock * (kWave / NMAX);
    const int blocks = ceil_div(s.count, per_block);
    const int grid = s.dev_count ? persistent_grid(blocks, 8) : blocks;
    potrf_reg_kernel<NMAX, SOLVE, VAR><<<grid, kWave * kWavesPerBlock, 0, st>>>(s);
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

}  // namespace lu_reg_detail

// One translation unit per NMAX (lu_reg_<NMAX>.hip) instantiates these.
// var: per-matrix sizes in s.n_arr / s.m_arr / s.lda_arr (vbatched buckets).
template <int NMAX> magma_int_t launch_getrf_reg_n(const MatSet &s, bool solve, bool var, hipStream_t st);
template <int NMAX> magma_int_t launch_potrf_reg_n(const MatSet &s, bool solve, bool var, hipStream_t st);

}  // namespace magma_amd

#define MAGMA_AMD_LU_REG_INSTANTIATE(N)                                                                    \
    namespace magma_amd {                                                                                  \
    template <> magma_int_t launch_getrf_reg_n<N>(const MatSet &s, bool solve, bool var, hipStream_t st)   \
    {                                                                                                      \
        using namespace lu_reg_detail;                                                                     \
        if (var) return solve ? launch_getrf_reg<N, true, true>(s, st) : launch_getrf_reg<N, false, true>(s, st); \
        return solve ? launch_getrf_reg<N, true, false>(s, st) : launch_getrf_reg<N, false, false>(s, st); \
    }                                                                                                      \
    template <> magma_int_t launch_potrf_reg_n<N>(const MatSet &s, bool solve, bool var, hipStream_t st)   \
    {                                                                                                      \
        using namespace lu_reg_detail;                                                                     \
        if (var) return solve ? launch_potrf_reg<N, true, true>(s, st) : launch_potrf_reg<N, false, true>(s, st); \
        return solve ? launch_potrf_reg<N, true, false>(s, st) : launch_potrf_reg<N, false, false>(s, st); \
    }                                                                                                      \
    }
