// lu_blocked.hip -- blocked right-looking LU and Cholesky, one 256-thread
// workgroup per matrix, for sizes above the register-resident kernels.
//
// Per panel of PW columns:
//   panel    the whole workgroup factors rows j0..m-1 x PW columns held in
//            registers (RPT rows per thread); pivot search by wavefront
//            shuffles + an LDS combine; two barriers per column.
//   laswp    the panel's interchanges applied, in LAPACK order, to the other
//            columns (thread per column).
//   trsm     U12 = L11^{-1} A12, thread per column, L11 from LDS.
//   gemm     A22 -= L21 U12 in 64x64 tiles, 4x4 per thread, operands in LDS.
// Replaces the reference's recursive panel + laswp + trsm + gemm launch chain
// (src/zgetrf_batched.cpp:222-351, src/zgetrf_panel_batched.cpp:127-243) and
// its lg Cholesky (src/zpotrf_batched.cpp:20-221) with one launch.
//
// Every element still receives its updates one fma at a time in increasing k
// (the netlib dgetf2 / dpotrf2 order), so the factors are bit-identical to the
// CPU restatement regardless of PW.
#include <cfloat>
#include <climits>

#include "kernels.h"

namespace magma_amd {
namespace {

constexpr int kThreads = 256;
constexpr int kTile = 64;

__device__ __forceinline__ int pick(const int *arr, int fixed, int id) { return arr ? arr[id] : fixed; }

// (key, pos) order of the pivot search: larger key wins, ties to smaller pos.
__device__ __forceinline__ void pivot_combine(double &key, int &kp, double ok, int op)
{
    const bool take = (ok > key) || (ok == key && op < kp);
    key = take ? ok : key;
    kp = take ? op : kp;
}

// A22 -= L21 * U12 (or the lower triangle of A22 -= L21 * L21' when SYRK) for
// rows [r0, m) x cols [c0, n) of A, k = kb columns starting at column k0 (L21)
// and row k0 (U12).  Element (i,c) gets c = fma(-l_ik, u_kc, c) in increasing k.
template <bool SYRK>
__device__ void trailing_update(double *A, int lda, int r0, int m, int c0, int n, int k0, int kb,
                                bool exact, double (*Ls)[17], double (*Us)[kTile + 1])
{
    const int tid = threadIdx.x;
    const int tr = tid % 16;             // 4-row block within the tile
    const int tc = tid / 16;             // 4-col block within the tile
    for (int tcol = c0; tcol < n; tcol += kTile) {
        for (int trow = SYRK ? tcol : r0; trow < m; trow += kTile) {
            // stage L21 tile (kTile x kb) and U12 tile (kb x kTile)
            for (int e = tid; e < kTile * 16; e += kThreads) {
                const int i = e % kTile, k = e / kTile;
                const int gi = trow + i;
                Ls[i][k] = (k < kb && gi < m) ? A[gi + (size_t)(k0 + k) * lda] : 0.0;
            }
            for (int e = tid; e < kTile * 16; e += kThreads) {
                const int c = e / 16, k = e % 16;
                const int gc = tcol + c;
                double u = 0.0;
                if (k < kb && gc < n) u = SYRK ? A[gc + (size_t)(k0 + k) * lda] : A[(k0 + k) + (size_t)gc * lda];
                Us[k][c] = u;
            }
            __syncthreads();
            double acc[4][4];
#pragma unroll
            for (int jc = 0; jc < 4; ++jc) {
                const int gc = tcol + tc * 4 + jc;
#pragma unroll
                for (int ir = 0; ir < 4; ++ir) {
                    const int gi = trow + tr * 4 + ir;
                    const bool ok = gi < m && gc < n && (!SYRK || gi >= gc);
                    acc[ir][jc] = ok ? A[gi + (size_t)gc * lda] : 0.0;
                }
            }
            if (!exact) {
                for (int k = 0; k < kb; ++k) {
                    double l[4], u[4];
#pragma unroll
                    for (int ir = 0; ir < 4; ++ir) l[ir] = Ls[tr * 4 + ir][k];
#pragma unroll
                    for (int jc = 0; jc < 4; ++jc) u[jc] = -Us[k][tc * 4 + jc];
#pragma unroll
                    for (int ir = 0; ir < 4; ++ir)
#pragma unroll
                        for (int jc = 0; jc < 4; ++jc) acc[ir][jc] = fma(l[ir], u[jc], acc[ir][jc]);
                }
            } else {
                // netlib skips a zero pivot-row entry; differs only with Inf/NaN multipliers
                for (int k = 0; k < kb; ++k) {
#pragma unroll
                    for (int ir = 0; ir < 4; ++ir)
#pragma unroll
                        for (int jc = 0; jc < 4; ++jc) {
                            const double u = Us[k][tc * 4 + jc];
                            if (u != 0.0) acc[ir][jc] = fma(Ls[tr * 4 + ir][k], -u, acc[ir][jc]);
                        }
                }
            }
#pragma unroll
            for (int jc = 0; jc < 4; ++jc) {
                const int gc = tcol + tc * 4 + jc;
#pragma unroll
                for (int ir = 0; ir < 4; ++ir) {
                    const int gi = trow + tr * 4 + ir;
                    if (gi < m && gc < n && (!SYRK || gi >= gc)) A[gi + (size_t)gc * lda] = acc[ir][jc];
                }
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
template <int RPT, int PW>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) getrf_blocked_kernel(MatSet s)
{
    static_assert(PW <= 16, "panel width");
    __shared__ double prow[PW];
    __shared__ double red_key[kThreads / 64];
    __shared__ int red_pos[kThreads / 64];
    __shared__ int swaps[PW];
    __shared__ double L11[PW][PW + 1];
    __shared__ double Ls[kTile][17];
    __shared__ double Us[16][kTile + 1];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    const int cnt = item_count(s);
    const int *idx = item_index(s);
    for (int item = blockIdx.x; item < cnt; item += gridDim.x) {
        const int id = idx ? idx[item] : item;
        const int m = pick(s.m_arr, s.m, id);
        const int n = pick(s.n_arr, s.n, id);
        const int lda = pick(s.lda_arr, s.lda, id);
        double *A = s.A[id];
        int *ipiv = s.ipiv[id];
        const int mn = min(m, n);
        int info = 0;

        for (int j0 = 0; j0 < mn; j0 += PW) {
            const int jb = min(PW, mn - j0);
            const int mp = m - j0;
            for (int e = tid; e < PW * (PW + 1); e += kThreads) (&L11[0][0])[e] = 0.0;
            double a[RPT][PW];
            int pos[RPT];
    #pragma unroll
            for (int i = 0; i < RPT; ++i) {
                const int pr = tid + kThreads * i;
                pos[i] = pr;
    #pragma unroll
                for (int c = 0; c < PW; ++c)
                    a[i][c] = (pr < mp && c < jb) ? A[(j0 + pr) + (size_t)(j0 + c) * lda] : 0.0;
            }
            bool bad = false;
    #pragma unroll
            for (int jj = 0; jj < PW; ++jj) {
                if (jj < jb) {
                    double key = -1.0;
                    int kp = INT_MAX;
    #pragma unroll
                    for (int i = 0; i < RPT; ++i) {
                        const int pr = tid + kThreads * i;
                        const bool cand = pr < mp && pos[i] >= jj;
                        const double v = a[i][jj];
                        double k2 = -1.0;
                        if (cand) k2 = (v != v) ? ((pos[i] == jj) ? __builtin_huge_val() : -1.0) : fabs(v);
                        pivot_combine(key, kp, k2, cand ? pos[i] : INT_MAX);
                    }
    #pragma unroll
                    for (int off = 32; off > 0; off >>= 1)
                        pivot_combine(key, kp, __shfl_xor(key, off), __shfl_xor(kp, off));
                    if (lane == 0) {
                        red_key[w] = key;
                        red_pos[w] = kp;
                    }
                    __syncthreads();
                    key = red_key[0];
                    kp = red_pos[0];
    #pragma unroll
                    for (int q = 1; q < kThreads / 64; ++q) pivot_combine(key, kp, red_key[q], red_pos[q]);
                    const int p = kp;
    #pragma unroll
                    for (int i = 0; i < RPT; ++i) {
                        const int pr = tid + kThreads * i;
                        if (pr < mp && pos[i] == p) {
    #pragma unroll
                            for (int c = jj; c < PW; ++c) prow[c] = a[i][c];
                        }
                    }
                    if (tid == 0) swaps[jj] = p;
                    __syncthreads();
                    const double piv = prow[jj];
                    if (piv == 0.0 && info == 0) info = j0 + jj + 1;
                    const bool use_rcp = fabs(piv) >= DBL_MIN;
                    const double rc = 1.0 / piv;
    #pragma unroll
                    for (int i = 0; i < RPT; ++i) {
                        const int pr = tid + kThreads * i;
                        pos[i] = (pos[i] == jj) ? p : ((pos[i] == p) ? jj : pos[i]);
                        const bool below = pr < mp && pos[i] > jj;
                        if (below) {
                            double l = a[i][jj];
                            if (piv != 0.0) l = use_rcp ? rc * l : l / piv;
                            a[i][jj] = l;
                            bad |= !__builtin_isfinite(l);
    #pragma unroll
                            for (int c = jj + 1; c < PW; ++c) {
                                const double u = prow[c];
                                if (c < jb && u != 0.0) a[i][c] = fma(l, -u, a[i][c]);
                            }
                        }
                    }
                }
            }
            // panel back to global in pivoted order; L11 (unit lower) to LDS
    #pragma unroll
            for (int i = 0; i < RPT; ++i) {
                const int pr = tid + kThreads * i;
                if (pr < mp) {
    #pragma unroll
                    for (int c = 0; c < PW; ++c)
                        if (c < jb) A[(j0 + pos[i]) + (size_t)(j0 + c) * lda] = a[i][c];
                    if (pos[i] < jb) {
    #pragma unroll
                        for (int c = 0; c < PW; ++c) L11[pos[i]][c] = a[i][c];
                    }
                }
            }
            if (tid < jb) ipiv[j0 + tid] = j0 + swaps[tid] + 1;
            const bool exact = __syncthreads_or(bad);

            // laswp on the other columns, then U12 = L11^{-1} A12 on the right ones
            const int nother = n - jb;
            for (int t = tid; t < nother; t += kThreads) {
                const int c = (t < j0) ? t : t + jb;
                double *col = A + (size_t)c * lda;
                for (int jj = 0; jj < jb; ++jj) {
                    const int r2 = j0 + swaps[jj];
                    const int r1 = j0 + jj;
                    if (r2 != r1) {
                        const double x = col[r1];
                        col[r1] = col[r2];
                        col[r2] = x;
                    }
                }
                if (c >= j0 + jb) {
                    double x[PW];
    #pragma unroll
                    for (int k = 0; k < PW; ++k) x[k] = (k < jb) ? col[j0 + k] : 0.0;
                    // rows of L11 at and beyond jb are zero, so the last
                    // (narrower) panel needs no guard; one L11 column per
                    // step keeps the register footprint at the column itself
                    if (!exact) {
    #pragma unroll
                        for (int k = 0; k < PW; ++k) {
                            const double u = -x[k];
    #pragma unroll
                            for (int i = k + 1; i < PW; ++i) x[i] = fma(L11[i][k], u, x[i]);
                            __builtin_amdgcn_sched_barrier(0);
                        }
                    } else {
    #pragma unroll
                        for (int k = 0; k < PW; ++k) {
                            const double u = x[k];
    #pragma unroll
                            for (int i = k + 1; i < PW; ++i)
                                if (u != 0.0) x[i] = fma(L11[i][k], -u, x[i]);
                            __builtin_amdgcn_sched_barrier(0);
                        }
                    }
    #pragma unroll
                    for (int k = 0; k < PW; ++k)
                        if (k < jb) col[j0 + k] = x[k];
                }
            }
            __syncthreads();
            if (j0 + jb < m && j0 + jb < n)
                trailing_update<false>(A, lda, j0 + jb, m, j0 + jb, n, j0, jb, exact, Ls, Us);
        }
        if (tid == 0) s.info[id] = info;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
typedef double d4 __attribute__((ext_vector_type(4)));

// The lower triangle of A22 -= L21 L21' (rows and columns [c0, n), k = kb
// panel columns from k0) on v_mfma_f64_16x16x4f64, one 32 x 32 tile per
// wavefront at a time.  The product is formed transposed: MFMA row m is the
// tile column (A operand -L(col, k)), MFMA column the tile row (B operand
// L(row, k)), so lane (lk = lane / 16, li = lane % 16) holds, in register i of
// block (bi, bj), entry (row r0 + 16 bj + li, col q0 + 16 bi + lk + 4 i).
// The MFMA adds its four k terms to each entry as a sequential fma chain and
// the k slices come in increasing order, so every entry gets
// fma(-l_ck, l_ik, a) in increasing k as the netlib loop does (terms past kb
// are -0 x 0 and leave an entry as it is).
__device__ void syrk_update_mfma(double *A, int lda, int c0, int n, int k0, int kb)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int nt = (n - c0 + 31) / 32;
    // lower-triangle tiles (tr >= tc), column by column
    const int ntiles = nt * (nt + 1) / 2;
#pragma unroll 1
    for (int t = w; t < ntiles; t += kThreads / 64) {
        int tc = 0, rem = t;
        while (rem >= nt - tc) {
            rem -= nt - tc;
            ++tc;
        }
        const int tr = tc + rem;
        const int r0 = c0 + 32 * tr, q0 = c0 + 32 * tc;
        d4 acc[2][2];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = r0 + 16 * bj + li;
                    const int col = q0 + 16 * bi + lk + 4 * i;
                    acc[bi][bj][i] = (row < n && col < n && row >= col) ? A[row + (size_t)col * lda] : 0.0;
                }
#pragma unroll 1
        for (int kq = 0; kq < kb; kq += 4) {
            const int kk = kq + lk;
            double ua[2], lb[2];
#pragma unroll
            for (int bi = 0; bi < 2; ++bi) {
                const int col = q0 + 16 * bi + li;
                ua[bi] = (col < n && kk < kb) ? -A[col + (size_t)(k0 + kk) * lda] : -0.0;
            }
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                const int row = r0 + 16 * bj + li;
                lb[bj] = (row < n && kk < kb) ? A[row + (size_t)(k0 + kk) * lda] : 0.0;
            }
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
                    acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[bi], lb[bj], acc[bi][bj], 0, 0, 0);
        }
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = r0 + 16 * bj + li;
                    const int col = q0 + 16 * bi + lk + 4 * i;
                    if (row < n && col < n && row >= col) A[row + (size_t)col * lda] = acc[bi][bj][i];
                }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Cholesky, lower.  No pivoting: rows stay in place, the panel holds rows
// j0..n-1 of columns j0..j0+jb-1; the diagonal block's upper part is never
// written back.  The trailing update is the FP64 MFMA SYRK above (the VALU
// loops when a multiplier is not finite).
template <int RPT, int PW>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) potrf_blocked_kernel(MatSet s)
{
    __shared__ double lcol[PW];
    __shared__ double diag;
    __shared__ double Ls[kTile][17];
    __shared__ double Us[16][kTile + 1];

    const int tid = threadIdx.x;
    const int cnt = item_count(s);
    const int *idx = item_index(s);
    for (int item = blockIdx.x; item < cnt; item += gridDim.x) {
        const int id = idx ? idx[item] : item;
        const int n = pick(s.n_arr, s.n, id);
        const int lda = pick(s.lda_arr, s.lda, id);
        double *A = s.A[id];
        int info = 0;

        for (int j0 = 0; j0 < n && info == 0; j0 += PW) {
            const int jb = min(PW, n - j0);
            const int mp = n - j0;
            double a[RPT][PW];
    #pragma unroll
            for (int i = 0; i < RPT; ++i) {
                const int pr = tid + kThreads * i;
    #pragma unroll
                for (int c = 0; c < PW; ++c)
                    a[i][c] = (pr < mp && c < jb && c <= pr) ? A[(j0 + pr) + (size_t)(j0 + c) * lda] : 0.0;
            }
            bool bad = false;
    #pragma unroll
            for (int jj = 0; jj < PW; ++jj) {
                if (jj < jb && info == 0) {
                    if (tid == jj) diag = a[0][jj];
                    __syncthreads();
                    const double ajj = diag;
                    if (!(ajj > 0.0)) {
                        info = j0 + jj + 1;
                    } else {
                        const double ljj = sqrt(ajj);
                        const double rc = 1.0 / ljj;
                        if (tid == jj) a[0][jj] = ljj;
                        double l[RPT];
    #pragma unroll
                        for (int i = 0; i < RPT; ++i) {
                            const int pr = tid + kThreads * i;
                            l[i] = 0.0;
                            if (pr < mp && pr > jj) {
                                l[i] = rc * a[i][jj];
                                a[i][jj] = l[i];
                                bad |= !__builtin_isfinite(l[i]);
                                if (pr < jb) lcol[pr] = l[i];
                            }
                        }
                        __syncthreads();
    #pragma unroll
                        for (int i = 0; i < RPT; ++i) {
                            const int pr = tid + kThreads * i;
                            if (pr < mp && pr > jj) {
    #pragma unroll
                                for (int c = jj + 1; c < PW; ++c) {
                                    if (c < jb && c <= pr) {
                                        const double u = lcol[c];
                                        if (u != 0.0) a[i][c] = fma(-u, l[i], a[i][c]);
                                    }
                                }
                            }
                        }
                    }
                    __syncthreads();
                }
            }
    #pragma unroll
            for (int i = 0; i < RPT; ++i) {
                const int pr = tid + kThreads * i;
                if (pr < mp) {
    #pragma unroll
                    for (int c = 0; c < PW; ++c)
                        if (c < jb && c <= pr) A[(j0 + pr) + (size_t)(j0 + c) * lda] = a[i][c];
                }
            }
            const bool exact = __syncthreads_or(bad);
            if (info == 0 && j0 + jb < n) {
                // finite multipliers: FP64 MFMA; else the VALU loops with the
                // reference BLAS skip of zero entries (only then can it matter)
                if (!exact) syrk_update_mfma(A, lda, j0 + jb, n, j0, jb);
                else trailing_update<true>(A, lda, j0 + jb, n, j0 + jb, n, j0, jb, exact, Ls, Us);
            }
        }
        if (tid == 0) s.info[id] = info;
        __syncthreads();
    }
}

}  // namespace

magma_int_t launch_getrf_blocked(const MatSet &s, int max_m, hipStream_t st)
{
    if (s.count <= 0) return MAGMA_SUCCESS;
    const int grid = s.dev_count ? persistent_grid(s.count, 4) : s.count;
    if (max_m <= 256)       getrf_blocked_kernel<1, 16><<<grid, kThreads, 0, st>>>(s);
    else if (max_m <= 512)  getrf_blocked_kernel<2, 16><<<grid, kThreads, 0, st>>>(s);
    else if (max_m <= 1024) getrf_blocked_kernel<4, 8><<<grid, kThreads, 0, st>>>(s);
    else if (max_m <= 2048) getrf_blocked_kernel<8, 4><<<grid, kThreads, 0, st>>>(s);
    else if (max_m <= 4096) getrf_blocked_kernel<16, 2><<<grid, kThreads, 0, st>>>(s);
    else if (max_m <= 8192) getrf_blocked_kernel<32, 1><<<grid, kThreads, 0, st>>>(s);
    else return MAGMA_ERR_NOT_SUPPORTED;
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

magma_int_t launch_potrf_blocked(const MatSet &s, int max_n, hipStream_t st)
{
    if (s.count <= 0) return MAGMA_SUCCESS;
    const int grid = s.dev_count ? persistent_grid(s.count, 4) : s.count;
    if (max_n <= 256)       potrf_blocked_kernel<1, 16><<<grid, kThreads, 0, st>>>(s);
    else if (max_n <= 512)  potrf_blocked_kernel<2, 16><<<grid, kThreads, 0, st>>>(s);
    else if (max_n <= 1024) potrf_blocked_kernel<4, 8><<<grid, kThreads, 0, st>>>(s);
    else if (max_n <= 2048) potrf_blocked_kernel<8, 4><<<grid, kThreads, 0, st>>>(s);
    else if (max_n <= 4096) potrf_blocked_kernel<16, 2><<<grid, kThreads, 0, st>>>(s);
    else if (max_n <= 8192) potrf_blocked_kernel<32, 1><<<grid, kThreads, 0, st>>>(s);
    else return MAGMA_ERR_NOT_SUPPORTED;
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

}  // namespace magma_amd
