// lu_mfma.hip -- blocked right-looking LU with partial pivoting, FP64, one
// 256-thread workgroup (4 wavefronts) per matrix, for 64 < max(m, n) <= 512.
//
// Per panel of PW = 32 columns (rows j0..m-1), three phases separated by
// workgroup barriers:
//
//   panel  The panel lives in registers, row pr = tid + 256 i of it in thread
//          tid.  Per column: local (|a|, position) key, wavefront arg-max by
//          DPP (quad, half-row and row mirrors) plus a read of the four row
//          leaders, then each wavefront's winning row is published in an LDS
//          slot (double-buffered by column parity) and ONE barrier later every
//          thread picks the winning wavefront's row as the pivot row.  Row
//          interchanges inside the panel are logical (pos); rows are written
//          back to their final positions once, at the end of the panel.
//   rows   The panel's interchanges as a gather of the 32-row block right of
//          the panel into LDS, 32 lanes per column (coalesced), the old top
//          rows stored to the rows they move to; then one thread per column
//          solves U12 = L11^-1 A12 in registers (L11 from LDS) and leaves -U12
//          in LDS for the update; then U12 is stored back, 32 lanes per
//          column.  The columns left of the panel (the L part) take all later
//          interchanges in one pass after the last panel, a column read and
//          written once, instead of a gather per panel.
//   update A22 -= L21 U12 in 32x32 tiles, one tile per wavefront at a time, on
//          v_mfma_f64_16x16x4f64.  The product is formed transposed
//          (A22^T -= U12^T L21^T) so the C tile and the L21 operand are read
//          as contiguous 128-byte column segments; -U12^T comes from LDS.
//
// Replaces the reference's launch chain for these sizes (panel + laswp + trsm
// + gemm, src/zgetrf_batched.cpp:222-351, src/zgetrf_panel_batched.cpp:127-243,
// magmablas/zgetf2_kernels.cu:353-760) with one persistent launch.
//
// Rounding: every entry receives its updates as single fused multiply-adds
// in increasing k, in the panel (VALU), the solve (VALU) and the update: the
// FP64 MFMA applies its four k terms of a slice as the sequential fma chain,
// slices in increasing k.  Multipliers are a * (1/pivot) (division below
// DBL_MIN), the pivot is the first maximum of |a| with NaN taken only in the
// first position.  So factors and pivots are the bits of netlib dgetrf with
// that update order.  When a multiplier of the panel is not finite the
// update and solve switch to VALU loops that skip zero U entries as the
// reference BLAS loops do (only then can the skip change a result).
#include <cfloat>
#include <climits>

#include "kernels.h"
#include "tune.h"
#include "wave_reduce.h"

namespace magma_amd {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int PW = 32;

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int pick(const int *arr, int fixed, int id) { return arr ? arr[id] : fixed; }

struct alignas(16) PivSlot {
    unsigned long long key;   // pivot_key of the wavefront's best row
    int pos;
    int pad;
    double row[PW];           // 16-byte aligned: written in pairs
};

template <int MT>
struct LuSmem {
    PivSlot slot[2][kWaves];
    double L11[PW][PW + 1];
    double Us[PW][MT + 1];    // -U12 of the current panel, [k][column - (j0 + jb)]; the odd
                              // pitch spreads a column's entries over the banks
    int src_top[PW];          // logical panel row r came from physical panel row src_top[r]
    int down_dst[PW];         // physical panel row r (< jb) moves to logical row down_dst[r] (>= jb), or -1
    int swaps[PW];
    int pst[MT / PW][PW];     // src_top of every panel, for the deferred left-column interchanges
    int pdd[MT / PW][PW];     // down_dst of every panel
};

template <int MT>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(MT <= 256 ? 2 : 1))) getrf_mfma_kernel(MatSet s)
{
    constexpr int RPT = MT / kThreads;
    __shared__ LuSmem<MT> sm;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    const int li = lane & 15;
    const int lk = lane >> 4;
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
            // ------------------------------------------------ panel --------
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
                    // ordered integer keys: |a| + 1 ulp of the bit pattern, 0 for
                    // rows that are not candidates and for NaN past the first
                    // position (idamax never takes those)
                    unsigned long long key = 0;
                    int kp = INT_MAX;
                    int own = 0;
#pragma unroll
                    for (int i = 0; i < RPT; ++i) {
                        const int pr = tid + kThreads * i;
                        const bool cand = pr < mp && pos[i] >= jj;
                        const double v = a[i][jj];
                        unsigned long long k2 = 0;
                        if (cand && (v == v || pos[i] == jj)) k2 = pivot_key(v == v ? fabs(v) : __builtin_huge_val());
                        const int p2 = cand ? pos[i] : INT_MAX;
                        if ((k2 > key) || (k2 == key && p2 < kp)) {
                            key = k2;
                            kp = p2;
                            own = i;
                        }
                    }
                    unsigned long long wk = key;
                    int wp = kp;
                    wave_argmax_key(wk, wp);
                    PivSlot &sl = sm.slot[jj & 1][w];
                    if (wp != INT_MAX && kp == wp) {
                        // 16-byte stores from the pair holding column jj on
                        // (the entries left of jj are written too, and unread)
                        double2 *dst = reinterpret_cast<double2 *>(sl.row);
#pragma unroll
                        for (int c = jj & ~1; c < PW; c += 2) {
                            double v0 = a[0][c], v1 = a[0][c + 1];
#pragma unroll
                            for (int i = 1; i < RPT; ++i) {
                                v0 = (own == i) ? a[i][c] : v0;
                                v1 = (own == i) ? a[i][c + 1] : v1;
                            }
                            dst[c / 2] = make_double2(v0, v1);
                        }
                    }
                    if (lane == 0) {
                        sl.key = wk;
                        sl.pos = wp;
                    }
                    __syncthreads();
                    // all four candidates, then the winner's row: two LDS
                    // round trips, each issued as one batch of reads
                    unsigned long long ck[kWaves];
                    int cp[kWaves];
#pragma unroll
                    for (int q = 0; q < kWaves; ++q) {
                        ck[q] = sm.slot[jj & 1][q].key;
                        cp[q] = sm.slot[jj & 1][q].pos;
                    }
                    asm volatile("" ::: "memory");
                    unsigned long long bk = ck[0];
                    int bp = cp[0];
                    int bw = 0;
#pragma unroll
                    for (int q = 1; q < kWaves; ++q) {
                        const bool take = (ck[q] > bk) | ((ck[q] == bk) & (cp[q] < bp));
                        bk = take ? ck[q] : bk;
                        bp = take ? cp[q] : bp;
                        bw = take ? q : bw;
                    }
                    const double *prow = sm.slot[jj & 1][bw].row;
                    // one batch of reads into registers when they fit beside
                    // the panel (one panel row per thread); else read per use
                    double u[PW];
#pragma unroll
                    for (int c = jj & ~1; c < PW; c += 2) {
                        const double2 t = reinterpret_cast<const double2 *>(prow)[c / 2];
                        u[c] = t.x;
                        u[c + 1] = t.y;
                    }
                    if (RPT == 1) asm volatile("" ::: "memory");
                    const int p = bp;
                    if (tid == 0) sm.swaps[jj] = p;
                    const double piv = u[jj];
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
                            const bool fin = __builtin_isfinite(l);
                            bad |= !fin;
                            if (__builtin_amdgcn_ballot_w64(!fin) == 0) {
                                // finite multipliers: a zero pivot-row entry adds
                                // a zero, so the reference dger's skip of it
                                // changes nothing (up to the sign of a zero)
#pragma unroll
                                for (int c = jj + 1; c < PW; ++c) a[i][c] = fma(l, -u[c], a[i][c]);
                            } else {
                                // reference dger: a zero pivot-row entry is
                                // skipped.  fma(0, -0, x) is x exactly, so a zero
                                // multiplier there leaves the entry as it was.
                                double ls = l;
                                asm volatile("" : "+v"(ls));      // keeps the two loops apart
#pragma unroll
                                for (int c = jj + 1; c < PW; ++c)
                                    a[i][c] = fma((u[c] != 0.0) ? ls : 0.0, -u[c], a[i][c]);
                            }
                        }
                    }
                }
            }
            // panel rows to their final positions; L11, the interchange maps
#pragma unroll
            for (int i = 0; i < RPT; ++i) {
                const int pr = tid + kThreads * i;
                if (pr < mp) {
#pragma unroll
                    for (int c = 0; c < PW; ++c)
                        if (c < jb) A[(j0 + pos[i]) + (size_t)(j0 + c) * lda] = a[i][c];
                    if (pos[i] < jb) {
#pragma unroll
                        for (int c = 0; c < PW; ++c) sm.L11[pos[i]][c] = a[i][c];
                        sm.src_top[pos[i]] = pr;
                    }
                    if (pr < jb) sm.down_dst[pr] = (pos[i] >= jb) ? pos[i] : -1;
                }
            }
            if (tid < jb) ipiv[j0 + tid] = j0 + sm.swaps[tid] + 1;
            const bool exact = __syncthreads_or(bad);
            // src_top entries come from the threads owning the moved rows, in
            // any wavefront: read them after the barrier
            if (tid < jb) {
                sm.pst[j0 / PW][tid] = sm.src_top[tid];
                sm.pdd[j0 / PW][tid] = sm.down_dst[tid];
            }

            // ------------------------------------------------ rows ---------
            const int cb = j0 + jb;                 // first column right of the panel
            // Columns right of the panel only: the columns left of it (the L
            // part) get all later interchanges in one pass after the last
            // panel.  Three passes over the jb x nright block U12:
            //   gather  32 consecutive lanes per column (coalesced 256-byte
            //           reads): logical top row r from panel row src_top[r]
            //           into Us[r][column], the old top row r to down_dst[r];
            //           a column's reads all precede its writes (one
            //           half-wavefront owns it);
            //   solve   thread per column, U12 = L11^-1 A12 on Us in place
            //           (stored negated, the update's operand);
            //   store   32 lanes per column again, U12 back to A.
            const int nright = n - cb;
            constexpr int kCPI = kThreads / PW;     // columns per pass of the workgroup
            constexpr int kBatch = 8;               // passes whose loads are in flight together
            {
                const int r = tid & (PW - 1);
                const int rs = min(r, jb - 1);
                const int src = sm.src_top[rs];
                const int dd = (r < jb) ? sm.down_dst[r] : -1;
#pragma unroll 1
                for (int cl0 = tid / PW; cl0 < nright; cl0 += kCPI * kBatch) {
                    double xv[kBatch], wv[kBatch];
#pragma unroll
                    for (int u = 0; u < kBatch; ++u) {
                        const double *col = A + (size_t)(cb + min(cl0 + kCPI * u, nright - 1)) * lda + j0;
                        xv[u] = col[src];
                        wv[u] = col[rs];
                    }
#pragma unroll
                    for (int u = 0; u < kBatch; ++u) {
                        const int cl = cl0 + kCPI * u;
                        if (cl < nright) {
                            if (dd >= 0) A[(j0 + dd) + (size_t)(cb + cl) * lda] = wv[u];
                            sm.Us[r][cl] = (r < jb) ? xv[u] : 0.0;
                        }
                    }
                }
            }
            __syncthreads();
#pragma unroll 1
            for (int cl = tid; cl < nright; cl += kThreads) {
                double x[PW];
#pragma unroll
                for (int k = 0; k < PW; ++k) x[k] = sm.Us[k][cl];
                // rows of L11 at and beyond jb are never read for k < jb
                // into a stored result: x[k] = 0 there
                if (!exact) {
#pragma unroll
                    for (int k = 0; k < PW; ++k) {
                        const double u = -x[k];
#pragma unroll
                        for (int i = k + 1; i < PW; ++i) x[i] = fma(sm.L11[i][k], u, x[i]);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < PW; ++k) {
                        const double u = x[k];
#pragma unroll
                        for (int i = k + 1; i < PW; ++i)
                            if (u != 0.0) x[i] = fma(sm.L11[i][k], -u, x[i]);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
#pragma unroll
                for (int k = 0; k < PW; ++k) sm.Us[k][cl] = (k < jb) ? -x[k] : -0.0;
            }
            __syncthreads();
            {
                const int r = tid & (PW - 1);
                if (r < jb) {
#pragma unroll 1
                    for (int cl0 = tid / PW; cl0 < nright; cl0 += kCPI * kBatch) {
                        double v[kBatch];
#pragma unroll
                        for (int u = 0; u < kBatch; ++u) v[u] = sm.Us[r][min(cl0 + kCPI * u, nright - 1)];
#pragma unroll
                        for (int u = 0; u < kBatch; ++u)
                            if (cl0 + kCPI * u < nright) A[(j0 + r) + (size_t)(cb + cl0 + kCPI * u) * lda] = -v[u];
                    }
                }
            }
            // no barrier: the update reads Us and rows >= cb only

            // ------------------------------------------------ update -------
            const int nrt = (m - cb + 31) / 32;
            const int nct = (n - cb + 31) / 32;
            if (m > cb && n > cb) {
#pragma unroll 1
                for (int t = w; t < nrt * nct; t += kWaves) {
                    const int r0 = cb + 32 * (t % nrt);
                    const int c0 = cb + 32 * (t / nrt);
                    if (!exact) {
                        // D = C^T: lane (lk, li), register i of block (bi, bj)
                        // holds C(r0 + 16 bj + li, c0 + 16 bi + lk + 4 i)
                        d4 acc[2][2];
#pragma unroll
                        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                                for (int i = 0; i < 4; ++i) {
                                    const int row = r0 + 16 * bj + li;
                                    const int col = c0 + 16 * bi + lk + 4 * i;
                                    acc[bi][bj][i] = (row < m && col < n) ? A[row + (size_t)col * lda] : 0.0;
                                }
#pragma unroll
                        for (int k0 = 0; k0 < PW; k0 += 4) {
                            const int kk = k0 + lk;
                            double ua[2], lb[2];
#pragma unroll
                            for (int bi = 0; bi < 2; ++bi) {
                                const int col = c0 + 16 * bi + li;
                                ua[bi] = (col < n) ? sm.Us[kk][col - cb] : -0.0;
                            }
#pragma unroll
                            for (int bj = 0; bj < 2; ++bj) {
                                const int row = r0 + 16 * bj + li;
                                lb[bj] = (row < m && kk < jb) ? A[row + (size_t)(j0 + kk) * lda] : 0.0;
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
                                    const int col = c0 + 16 * bi + lk + 4 * i;
                                    if (row < m && col < n) A[row + (size_t)col * lda] = acc[bi][bj][i];
                                }
                    } else {
                        // reference BLAS loop semantics: skip a zero U entry
                        for (int e = lane; e < 32 * 32; e += 64) {
                            const int row = r0 + (e & 31), col = c0 + (e >> 5);
                            if (row < m && col < n) {
                                double v = A[row + (size_t)col * lda];
                                for (int k = 0; k < jb; ++k) {
                                    const double u = -sm.Us[k][col - cb];
                                    if (u != 0.0) v = fma(A[row + (size_t)(j0 + k) * lda], -u, v);
                                }
                                A[row + (size_t)col * lda] = v;
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
        // Deferred interchanges of the L part: the columns of panel k get the
        // interchanges of panels k+1, ... in order (each as the panel applied
        // it: logical top row r from row st[r], old top row r to dd[r]),
        // through LDS (the U12 staging buffer), with one read and one write
        // of each column.
        {
            const int npan = (mn + PW - 1) / PW;
            double *buf = &sm.Us[0][0];
#pragma unroll 1
            for (int k = 0; k + 1 < npan; ++k) {
                const int c0 = PW * k, nc = min(PW, mn - c0);
                const int r0 = PW * (k + 1), nr = m - r0;
                __syncthreads();
                for (int e = tid; e < nc * nr; e += kThreads) {
                    const int cl = e / nr, r = e % nr;
                    buf[cl * MT + r] = A[(r0 + r) + (size_t)(c0 + cl) * lda];
                }
#pragma unroll 1
                for (int p = k + 1; p < npan; ++p) {
                    const int off = PW * p - r0, jbp = min(PW, mn - PW * p);
                    __syncthreads();
                    constexpr int TPT = PW * PW / kThreads;
                    double sv[TPT], tv[TPT];
                    int base[TPT], rr[TPT], dd[TPT];     // base: column cl's row off
#pragma unroll
                    for (int u = 0; u < TPT; ++u) {
                        const int e = tid + kThreads * u;
                        const int cl = e / PW, r = e % PW;
                        const bool on = cl < nc && r < jbp;
                        base[u] = on ? cl * MT + off : -1;
                        rr[u] = r;
                        dd[u] = on ? sm.pdd[p][r] : -1;
                        if (on) {
                            sv[u] = buf[base[u] + sm.pst[p][r]];
                            tv[u] = buf[base[u] + r];
                        }
                    }
                    __syncthreads();
#pragma unroll
                    for (int u = 0; u < TPT; ++u) {
                        if (base[u] >= 0) {
                            buf[base[u] + rr[u]] = sv[u];
                            if (dd[u] >= 0) buf[base[u] + dd[u]] = tv[u];
                        }
                    }
                }
                __syncthreads();
                for (int e = tid; e < nc * nr; e += kThreads) {
                    const int cl = e / nr, r = e % nr;
                    A[(r0 + r) + (size_t)(c0 + cl) * lda] = buf[cl * MT + r];
                }
            }
        }
        if (tid == 0) s.info[id] = info;
    }
}

}  // namespace

template <class K>
static int resident_per_cu(K kernel)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kThreads, 0) != hipSuccess || nb < 1) nb = 1;
    return nb;
}

magma_int_t launch_getrf_blocked(const MatSet &s, int max_m, hipStream_t st);
magma_int_t launch_getrf_reg128(const MatSet &s, int max_dim, int waves_per_simd, hipStream_t st);

// 64 < max(m, n) <= 512; returns MAGMA_ERR_NOT_SUPPORTED above.  Variant
// (tuned per size class): 0 as many workgroups per CU as fit, 1 one workgroup
// per CU, 2 the VALU blocked kernel, 3 / 4 the register-resident kernel
// (lu_reg128.hip) at three / two wavefronts per SIMD, for max(m, n) <= 128
// (variant 0 above).
magma_int_t launch_getrf_mfma(const MatSet &s, int max_dim, hipStream_t st)
{
    if (s.count <= 0) return MAGMA_SUCCESS;
    if (max_dim > 512) return MAGMA_ERR_NOT_SUPPORTED;
    int v = tuned_variant(kTuneGetrf, size_class_of(max_dim));
    if (v >= 3 && max_dim <= 128) return launch_getrf_reg128(s, max_dim, v == 3 ? 3 : 2, st);
    if (v >= 3) v = 0;
    if (v == 2) return launch_getrf_blocked(s, max_dim, st);
    if (max_dim <= 256) {
        const int grid = persistent_grid(s.count, v == 1 ? 1 : resident_per_cu(getrf_mfma_kernel<256>));
        getrf_mfma_kernel<256><<<grid, kThreads, 0, st>>>(s);
    } else {
        const int grid = persistent_grid(s.count, v == 1 ? 1 : resident_per_cu(getrf_mfma_kernel<512>));
        getrf_mfma_kernel<512><<<grid, kThreads, 0, st>>>(s);
    }
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

}  // namespace magma_amd
