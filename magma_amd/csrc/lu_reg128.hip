// lu_reg128.hip -- register-resident LU with partial pivoting for
// 64 < max(m, n) <= 128: one 256-thread workgroup per matrix, the whole
// matrix in VGPRs for the whole factorization (no pass through memory between
// columns).
//
// Wavefront w holds columns [32 w, 32 w + 32); lane l holds physical rows l
// and l + 64 of them (64 doubles, four 16-wide vectors).  Every wavefront
// holds the same rows in the same lanes, so each keeps its own copy of the
// row bookkeeping (logical positions, "below the pivot" masks) and the only
// data that crosses wavefronts is the pivot column: per column j one LDS
// record (pivot position, pivot value, the finished column j by physical row)
// and ONE workgroup barrier.
//
// Step j, every wavefront: read record j, apply the interchange to the
// logical positions, then the rank-1 update a(i, c) = fma(l_i, -u_c, a(i, c))
// of its columns c > j for the rows below the pivot, with u_c broadcast from
// the pivot row's lane through a per-wavefront LDS row (one lane writes, all
// read the same address).  Look-ahead: the wavefront that holds column
// j + 1 first updates that one column, selects its pivot (first maximum of
// |a| over the rows below, by DPP) and publishes record j + 1 before its
// other columns, so the next step's pivot search is off the critical path of
// the other wavefronts.  The holder reads that column by a uniform-index
// register move (s_set_gpr_idx) and stores the finished column to A in
// physical row order; the pivot columns are put in logical row order at the
// end (the matrix was loaded whole at the start, so A is free until then).
//
// Finished columns (c <= j) are not branched around: their u_c is zero.  With
// finite multipliers that adds a signed zero (the value is unchanged); when
// any multiplier of the step is Inf or NaN the update switches to the exact
// form, which skips a zero u_c as the reference dger loop does.
//
// Numerics are netlib dgetf2's (every element updated by one fma per k in
// increasing k, multipliers a * (1/pivot), division below DBL_MIN, idamax's
// first maximum with NaN taken only in the first position), so factors and
// pivots equal those of the CPU restatement that the tests compare against.
// Replaces, for these sizes, the reference's panel + laswp + trsm + gemm launch
// chain (src/zgetrf_batched.cpp:222-351, src/zgetrf_panel_batched.cpp:127-243).
#include <cfloat>
#include <climits>

#include "kernels.h"
#include "lu_reg.h"
#include "wave_reduce.h"

namespace magma_amd {
namespace {

using lu_reg_detail::readlane_f64;

constexpr int kThreads = 256;
constexpr int kCols = 32;          // columns per wavefront

typedef double d16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ int pick(const int *arr, int fixed, int id) { return arr ? arr[id] : fixed; }

// column c (wave-uniform, 0..31) of one row slot
__device__ __forceinline__ double col_get(const d16 (&v)[2], int c) { return c < 16 ? v[0][c] : v[1][c - 16]; }

// Pivot of column jn over the candidate rows (logical position >= jn) from the
// column's values x0, x1 (rows lane, lane + 64); publishes the record of step
// jn (position, value, finished column) into slot pb and returns the finished
// column values of this lane's rows in x0, x1.
__device__ __forceinline__ void select_pivot(double &x0, double &x1, int pos0, int pos1, bool ok0, bool ok1, int jn,
                                             int lane, int *pivpos, double *pivval, double *col, int *ipiv)
{
    const bool c0 = ok0 && pos0 >= jn, c1 = ok1 && pos1 >= jn;
    // ordered integer keys; NaN counts as the largest value at the first
    // candidate position and is never taken elsewhere (idamax)
    unsigned long long key = 0;
    int kp = INT_MAX;
    if (c0) {
        if (x0 == x0 || pos0 == jn) key = pivot_key(x0 == x0 ? fabs(x0) : __builtin_huge_val());
        kp = pos0;
    }
    if (c1) {
        const unsigned long long k1 =
            (x1 == x1 || pos1 == jn) ? pivot_key(x1 == x1 ? fabs(x1) : __builtin_huge_val()) : 0ull;
        if ((k1 > key) | ((k1 == key) & (pos1 < kp))) {
            key = k1;
            kp = pos1;
        }
    }
    wave_argmax_key(key, kp);
    const unsigned long long w0 = __builtin_amdgcn_ballot_w64(c0 && pos0 == kp);
    const unsigned long long w1 = __builtin_amdgcn_ballot_w64(c1 && pos1 == kp);
    const double src = w0 ? x0 : x1;
    const double piv = readlane_f64(src, (int)__builtin_ctzll(w0 ? w0 : w1));
    // rows below the pivot after this step's interchange: candidates but the pivot row
    const bool b0 = c0 && pos0 != kp, b1 = c1 && pos1 != kp;
    // piv is wave-uniform: one branch, one reciprocal on the usual path
    if (fabs(piv) >= DBL_MIN) {
        const double rc = 1.0 / piv;
        if (b0) x0 = rc * x0;
        if (b1) x1 = rc * x1;
    } else if (piv != 0.0) {
        if (b0) x0 = x0 / piv;
        if (b1) x1 = x1 / piv;
    }
    col[lane] = x0;
    col[lane + 64] = x1;
    if (lane == 0) {
        *pivpos = kp;
        *pivval = piv;
        if (ipiv) ipiv[jn] = kp + 1;
    }
}

// Rank-1 update of this wavefront's columns for step j with finite
// multipliers: u_c from the pivot row staged in LDS (urow, zero for finished
// columns), rows of slot s updated where bs.  Not branched around finished
// chunks: a uniform branch per chunk made the register allocator copy the
// matrix.
__device__ __forceinline__ void step_update(d16 (&a0)[2], d16 (&a1)[2], const double *urow, double l0, double l1,
                                            bool b0, bool b1)
{
#pragma unroll
    for (int q = 0; q < kCols / 8; ++q) {
        const int h = q >> 1, o = (q & 1) * 8;
        double u[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) u[k] = urow[8 * q + k];
        if (b0) {
#pragma unroll
            for (int k = 0; k < 8; ++k) a0[h][o + k] = fma(l0, -u[k], a0[h][o + k]);
        }
        if (b1) {
#pragma unroll
            for (int k = 0; k < 8; ++k) a1[h][o + k] = fma(l1, -u[k], a1[h][o + k]);
        }
    }
}

// Columns [cb, ce) of A, held by this wavefront and stored in physical row
// order (row r of lane r % 64), rewritten in logical order (row r to pos).
// Only this wavefront touches these columns, and a column is read whole
// before any of it is written.
__device__ void to_logical_rows(double *A, int lda, int cb, int ce, int lane, int pos0, int pos1, bool ok0, bool ok1)
{
    for (int c = cb; c < ce; ++c) {
        const double v0 = ok0 ? A[lane + (size_t)c * lda] : 0.0, v1 = ok1 ? A[lane + 64 + (size_t)c * lda] : 0.0;
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        if (ok0) A[pos0 + (size_t)c * lda] = v0;
        if (ok1) A[pos1 + (size_t)c * lda] = v1;
    }
}

// The rest of the factorization in memory, from step j on (j's record already
// applied to pos0/pos1, its multipliers l in column j of A), for a matrix
// whose multipliers are not all finite: the same steps with netlib's skip of
// a zero u_c, which the register path's zero-u treatment of finished columns
// cannot give.  Each wavefront keeps its own columns (and lanes their own
// rows), so global memory is only read back by the lanes and wavefront that
// wrote it; the pivot record still passes through LDS.  Ends with the rows
// moved to their logical positions.
__device__ void finish_exact(double *A, int lda, int m, int n, int mn, int j, int c0, int lane, int &pos0, int &pos1,
                             bool ok0, bool ok1, int &info, double (*col)[128], int *pivpos, double *pivval, int *ipiv)
{
    const int r0 = lane, r1 = lane + 64;
    for (int t = j; t < mn; ++t) {
        const int pb = t & 1;
        if (t > j) {
            const int kp = pivpos[pb];
            if (pivval[pb] == 0.0 && info == 0) info = t + 1;
            pos0 = (pos0 == t) ? kp : ((pos0 == kp) ? t : pos0);
            pos1 = (pos1 == t) ? kp : ((pos1 == kp) ? t : pos1);
        }
        const bool b0 = ok0 && pos0 > t, b1 = ok1 && pos1 > t;
        const double l0 = col[pb][r0], l1 = col[pb][r1];
        const unsigned long long p0 = __builtin_amdgcn_ballot_w64(ok0 && pos0 == t);
        const unsigned long long p1 = __builtin_amdgcn_ballot_w64(ok1 && pos1 == t);
        const int prow = p0 ? (int)__builtin_ctzll(p0) : 64 + (int)__builtin_ctzll(p1);
        for (int c = max(t + 1, c0); c < min(c0 + kCols, n); ++c) {
            const double u = A[prow + (size_t)c * lda];
            if (u != 0.0) {
                if (b0) A[r0 + (size_t)c * lda] = fma(l0, -u, A[r0 + (size_t)c * lda]);
                if (b1) A[r1 + (size_t)c * lda] = fma(l1, -u, A[r1 + (size_t)c * lda]);
            }
        }
        const int tn = t + 1;
        if (tn < mn && tn >= c0 && tn < c0 + kCols) {
            double x0 = ok0 ? A[r0 + (size_t)tn * lda] : 0.0, x1 = ok1 ? A[r1 + (size_t)tn * lda] : 0.0;
            select_pivot(x0, x1, pos0, pos1, ok0, ok1, tn, lane, &pivpos[pb ^ 1], &pivval[pb ^ 1], col[pb ^ 1], ipiv);
            if (ok0) A[r0 + (size_t)tn * lda] = x0;
            if (ok1) A[r1 + (size_t)tn * lda] = x1;
        }
        __syncthreads();
    }
    to_logical_rows(A, lda, c0, min(c0 + kCols, n), lane, pos0, pos1, ok0, ok1);
}

template <int WPE>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WPE))) getrf_reg128_kernel(MatSet s)
{
    __shared__ double col[2][128];        // finished pivot column by physical row, by step parity
    __shared__ double prow[4][kCols];     // each wavefront's part of the pivot row
    __shared__ double pivval[2];
    __shared__ int pivpos[2];

    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c0 = kCols * w;             // first column of this wavefront
    const int cnt = item_count(s);
    const int *idx = item_index(s);
    for (int item = blockIdx.x; item < cnt; item += gridDim.x) {
        const int id = idx ? idx[item] : item;
        const int m = pick(s.m_arr, s.m, id), n = pick(s.n_arr, s.n, id), lda = pick(s.lda_arr, s.lda, id);
        double *A = (double *)((GDoublePtr *)s.A)[id];
        int *ipiv = s.ipiv ? (int *)((GIntPtr *)s.ipiv)[id] : nullptr;
        const int mn = min(m, n);
        const bool ok0 = lane < m, ok1 = lane + 64 < m;

        d16 a0[2], a1[2];                 // rows lane and lane + 64, columns c0 .. c0 + 31
        if (mn > 0) {
            // clamped addresses, zero outside the matrix: no branch per load
            const size_t r0 = min(lane, m - 1), r1 = min(lane + 64, m - 1);
#pragma unroll
            for (int c = 0; c < kCols; ++c) {
                const bool cin = c0 + c < n;
                const size_t off = (size_t)min(c0 + c, n - 1) * lda;
                const double v0 = A[r0 + off], v1 = A[r1 + off];
                a0[c >> 4][c & 15] = (ok0 && cin) ? v0 : 0.0;
                a1[c >> 4][c & 15] = (ok1 && cin) ? v1 : 0.0;
            }
        }
        int pos0 = lane, pos1 = lane + 64;        // logical positions (rows >= m never move)
        int info = 0;
        int exact_from = -1;                      // step at which the exact path took over

        if (mn > 0) {
            if (w == 0) {
                double x0 = a0[0][0], x1 = a1[0][0];
                select_pivot(x0, x1, pos0, pos1, ok0, ok1, 0, lane, &pivpos[0], &pivval[0], col[0], ipiv);
                if (ok0) A[lane] = x0;
                if (ok1) A[lane + 64] = x1;
            }
            __syncthreads();
        }
#pragma unroll 1
        for (int j = 0; j < mn; ++j) {
            const int pb = j & 1;
            const int kp = pivpos[pb];
            const double pv = pivval[pb];
            const double l0 = col[pb][lane], l1 = col[pb][lane + 64];
            if (pv == 0.0 && info == 0) info = j + 1;
            pos0 = (pos0 == j) ? kp : ((pos0 == kp) ? j : pos0);
            pos1 = (pos1 == j) ? kp : ((pos1 == kp) ? j : pos1);
            const bool b0 = ok0 && pos0 > j, b1 = ok1 && pos1 > j;
            // every wavefront holds the same rows and reads the same record, so
            // this test agrees across the workgroup
            if (__builtin_amdgcn_ballot_w64((b0 && !__builtin_isfinite(l0)) || (b1 && !__builtin_isfinite(l1)))) {
                exact_from = j;
                break;
            }
            const int jn = j + 1;
            if (c0 + kCols > jn && c0 < n) {                // this wavefront has live columns
                const bool in0 = ok0 && pos0 == j, in1 = ok1 && pos1 == j;
                if (jn < mn && jn >= c0) {
                    // look-ahead (this wavefront holds the next pivot column):
                    // column jn updated first, its pivot published, the finished
                    // column stored to A in physical row order (registers keep
                    // no copy; the rows are put in order at the end)
                    const int cj = jn - c0;
                    double x0 = col_get(a0, cj), x1 = col_get(a1, cj);
                    const unsigned long long p0 = __builtin_amdgcn_ballot_w64(in0);
                    const unsigned long long p1 = __builtin_amdgcn_ballot_w64(in1);
                    const double u = readlane_f64(p0 ? x0 : x1, (int)__builtin_ctzll(p0 ? p0 : p1));
                    if (b0) x0 = fma(l0, -u, x0);
                    if (b1) x1 = fma(l1, -u, x1);
                    select_pivot(x0, x1, pos0, pos1, ok0, ok1, jn, lane, &pivpos[pb ^ 1], &pivval[pb ^ 1], col[pb ^ 1],
                                 ipiv);
                    if (ok0) A[lane + (size_t)jn * lda] = x0;
                    if (ok1) A[lane + 64 + (size_t)jn * lda] = x1;
                }
                // the pivot row's entries of this wavefront's columns into LDS,
                // finished columns (c <= j) as zero
                double *urow = prow[w];
                if (in0) {
#pragma unroll
                    for (int c = 0; c < kCols; ++c) urow[c] = a0[c >> 4][c & 15];
                }
                if (in1) {
#pragma unroll
                    for (int c = 0; c < kCols; ++c) urow[c] = a1[c >> 4][c & 15];
                }
                if (lane <= j - c0) urow[lane] = 0.0;
                lu_reg_detail::wave_lds_sync();
                step_update(a0, a1, urow, l0, l1, b0, b1);
            }
            __syncthreads();
        }

        if (exact_from < 0) {
            // columns past min(m, n) from registers; the pivot columns, stored
            // in physical row order when they finished, moved to logical order
#pragma unroll
            for (int c = 0; c < kCols; ++c) {
                if (c0 + c >= mn && c0 + c < n) {
                    if (ok0) A[pos0 + (size_t)(c0 + c) * lda] = a0[c >> 4][c & 15];
                    if (ok1) A[pos1 + (size_t)(c0 + c) * lda] = a1[c >> 4][c & 15];
                }
            }
            to_logical_rows(A, lda, c0, min(c0 + kCols, mn), lane, pos0, pos1, ok0, ok1);
        } else {
            // the live columns (c > j) join the finished ones in physical row
            // order, then the exact steps
#pragma unroll
            for (int c = 0; c < kCols; ++c) {
                if (c0 + c > exact_from && c0 + c < n) {
                    if (ok0) A[lane + (size_t)(c0 + c) * lda] = a0[c >> 4][c & 15];
                    if (ok1) A[lane + 64 + (size_t)(c0 + c) * lda] = a1[c >> 4][c & 15];
                }
            }
            finish_exact(A, lda, m, n, mn, exact_from, c0, lane, pos0, pos1, ok0, ok1, info, col, pivpos, pivval, ipiv);
        }
        if (tid == 0) s.info[id] = info;
    }
}

template <int WPE>
int reg128_per_cu()
{
    static int per_cu = 0;
    if (per_cu == 0) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, getrf_reg128_kernel<WPE>, kThreads, 0) != hipSuccess ||
            nb < 1)
            nb = 1;
        per_cu = nb;
    }
    return per_cu;
}

}  // namespace

// 64 < max(m, n) <= 128 (m, n <= 128); MAGMA_ERR_NOT_SUPPORTED above.
// waves_per_simd: 3 (registers capped at 168, a few spill) or 2 (no cap).
magma_int_t launch_getrf_reg128(const MatSet &s, int max_dim, int waves_per_simd, hipStream_t st)
{
    if (s.count <= 0) return MAGMA_SUCCESS;
    if (max_dim > 128) return MAGMA_ERR_NOT_SUPPORTED;
    if (waves_per_simd == 3)
        getrf_reg128_kernel<3><<<persistent_grid(s.count, reg128_per_cu<3>()), kThreads, 0, st>>>(s);
    else
        getrf_reg128_kernel<2><<<persistent_grid(s.count, reg128_per_cu<2>()), kThreads, 0, st>>>(s);
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

}  // namespace magma_amd
