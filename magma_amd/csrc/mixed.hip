// mixed.hip -- batched mixed-precision solve (dsgesv semantics, batched).
//
// Per matrix, one 256-thread workgroup (persistent over the batch):
//   1. ||A||_inf, max|a_ij| (FP64); W = fp32(A * 2^-e), 2^e >= max|a_ij| a
//      power of two, so every scaled entry is below 1 in magnitude and the FP16
//      operands below cannot overflow.
//   2. LU with partial pivoting of W in FP32, blocked by 32 columns: the panel
//      is factored in registers (VALU FP32), the trailing update
//      A22 -= L21 U12 runs on FP16 MFMA (v_mfma_f32_16x16x16f16) with each
//      operand split hi + lo (hi = fp16(x), lo = fp16(x - hi)) and three
//      products hi*hi + hi*lo + lo*hi accumulated in FP32, which keeps about 22
//      bits of each operand: plain FP16 or BF16 operands would leave the
//      refinement divergent for condition numbers above ~1e4.
//   3. x = U^-1 L^-1 P b in FP32; then the FP64 refinement loop of dsgesv
//      (src/zcgesv_gpu.cpp:143-365): r = b - A x in FP64, stop when
//      ||r_j||_inf <= ||x_j||_inf ||A||_inf eps sqrt(n) for every column j,
//      else d = LU32 \ r, x += d (FP64), at most MAGMA_MIXED_ITERMAX sweeps.
//   4. Matrices that do not converge (or whose FP32 factorization meets a zero
//      pivot) are appended to a device list; the plain FP64 getrf + getrs
//      kernels then run on exactly that list (no host round trip).
#include <cfloat>
#include <climits>

#include <algorithm>

#include "kernels.h"

namespace magma_amd {

namespace {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

constexpr int kT = 256;        // threads per workgroup
constexpr int kPW = 32;        // panel width
constexpr int kTile = 64;      // trailing-update tile (rows and columns)
constexpr int kMaxN = 512;

struct MixedArgs {
    int n, nrhs, lda, ldb, ldx;
    double **A, **B, **X;
    int **ipiv;
    int *info;
    int *iter;
    float *W;            // one n x n FP32 slot per workgroup
    int *fb_list;        // matrices that fall back to FP64
    int *fb_count;
    int count;
};

// max over the workgroup that propagates NaN (fmax would drop it, and a NaN
// residual must never pass the convergence test)
__device__ __forceinline__ double nanmax(double a, double b) { return (a > b || a != a) ? a : b; }

__device__ __forceinline__ double block_max(double v, double *red)
{
    for (int off = 32; off > 0; off >>= 1) v = nanmax(v, __shfl_xor(v, off));
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double r = red[0];
    for (int q = 1; q < kT / 64; ++q) r = nanmax(r, red[q]);
    return r;
}

__device__ __forceinline__ void pivot_combine(float &key, int &kp, float ok, int op)
{
    const bool take = (ok > key) || (ok == key && op < kp);
    key = take ? ok : key;
    kp = take ? op : kp;
}

struct Smem {
    // panel / swaps
    float prow[kPW];
    float red_key[kT / 64];
    int red_pos[kT / 64];
    int swaps[kPW];
    int src[kPW];
    int dst[kPW];
    float L11[kPW][kPW + 1];
    float T[kPW][kT];          // interchange staging, one column per thread
    // trailing update operands (fp16 hi/lo), rows/cols x k, k contiguous
    _Float16 Lh[kTile][kPW + 4], Ll[kTile][kPW + 4];
    _Float16 Uh[kTile][kPW + 4], Ul[kTile][kPW + 4];
    // solve / refinement vectors
    double x[kMaxN];
    double r[kMaxN];
    float y[kMaxN];
    int perm[kMaxN];
    double red[kT / 64];
    int flag;
};

// ---------------------------------------------------------------------------
// FP32 LU of W (n x n, ld n) with partial pivoting; W holds the LAPACK-layout
// factors on return, ipiv (1-based) in `ipiv`.  Returns the first zero pivot
// (1-based) or 0.
template <int RPT>
__device__ __forceinline__ int lu32(float *W, int n, int *ipiv, Smem &sm)
{
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    int info = 0;
    for (int j0 = 0; j0 < n; j0 += kPW) {
        const int jb = min(kPW, n - j0);
        const int mp = n - j0;
        for (int e = tid; e < kPW * (kPW + 1); e += kT) (&sm.L11[0][0])[e] = 0.0f;
        float a[RPT][kPW];
        int pos[RPT];
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
            const int pr = tid + kT * i;
            pos[i] = pr;
#pragma unroll
            for (int c = 0; c < kPW; ++c)
                a[i][c] = (pr < mp && c < jb) ? W[(j0 + pr) + (size_t)(j0 + c) * n] : 0.0f;
        }
        // ---- panel: pivot search by wave shuffles + LDS combine ----
#pragma unroll
        for (int jj = 0; jj < kPW; ++jj) {
            if (jj < jb) {
                float key = -1.0f;
                int kp = INT_MAX;
#pragma unroll
                for (int i = 0; i < RPT; ++i) {
                    const int pr = tid + kT * i;
                    const bool cand = pr < mp && pos[i] >= jj;
                    const float v = a[i][jj];
                    float k2 = -1.0f;
                    if (cand) k2 = (v != v) ? ((pos[i] == jj) ? __builtin_huge_valf() : -1.0f) : fabsf(v);
                    pivot_combine(key, kp, k2, cand ? pos[i] : INT_MAX);
                }
#pragma unroll
                for (int off = 32; off > 0; off >>= 1)
                    pivot_combine(key, kp, __shfl_xor(key, off), __shfl_xor(kp, off));
                if (lane == 0) {
                    sm.red_key[w] = key;
                    sm.red_pos[w] = kp;
                }
                __syncthreads();
                key = sm.red_key[0];
                kp = sm.red_pos[0];
#pragma unroll
                for (int q = 1; q < kT / 64; ++q) pivot_combine(key, kp, sm.red_key[q], sm.red_pos[q]);
                const int p = kp;
#pragma unroll
                for (int i = 0; i < RPT; ++i) {
                    const int pr = tid + kT * i;
                    if (pr < mp && pos[i] == p) {
#pragma unroll
                        for (int c = jj; c < kPW; ++c) sm.prow[c] = a[i][c];
                    }
                }
                if (tid == 0) sm.swaps[jj] = p;
                __syncthreads();
                const float piv = sm.prow[jj];
                if (piv == 0.0f && info == 0) info = j0 + jj + 1;
                const float rc = 1.0f / piv;
#pragma unroll
                for (int i = 0; i < RPT; ++i) {
                    const int pr = tid + kT * i;
                    pos[i] = (pos[i] == jj) ? p : ((pos[i] == p) ? jj : pos[i]);
                    if (pr < mp && pos[i] > jj && piv != 0.0f) {
                        const float l = a[i][jj] * rc;
                        a[i][jj] = l;
#pragma unroll
                        for (int c = jj + 1; c < kPW; ++c) a[i][c] = fmaf(l, -sm.prow[c], a[i][c]);
                    }
                }
            }
        }
        // ---- panel back in pivoted order; L11 and the row mapping to LDS ----
        if (tid < kPW) sm.dst[tid] = -1;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
            const int pr = tid + kT * i;
            if (pr < mp) {
#pragma unroll
                for (int c = 0; c < kPW; ++c)
                    if (c < jb) W[(j0 + pos[i]) + (size_t)(j0 + c) * n] = a[i][c];
                if (pos[i] < jb) {
                    sm.src[pos[i]] = pr;
#pragma unroll
                    for (int c = 0; c < kPW; ++c) sm.L11[pos[i]][c] = a[i][c];
                }
                if (pr < jb && pos[i] >= jb) sm.dst[pr] = pos[i];
            }
        }
        if (tid < jb) ipiv[j0 + tid] = j0 + sm.swaps[tid] + 1;
        __syncthreads();
        // ---- interchanges on the other columns (all loads first, then the
        //      stores), and U12 = L11^-1 A12 on the right ones ----
        for (int t0 = 0; t0 < n - jb; t0 += kT) {
            const int t = t0 + tid;
            const bool act = t < n - jb;
            const int c = (t < j0) ? t : t + jb;
            float *col = W + (size_t)c * n + j0;
            // old top rows of this column to LDS (one column per thread)
            if (act) {
#pragma unroll
                for (int k = 0; k < kPW; ++k)
                    if (k < jb) sm.T[k][tid] = col[k];
            }
            float nw[kPW];
#pragma unroll
            for (int k = 0; k < kPW; ++k) nw[k] = 0.0f;
            if (act) {
#pragma unroll
                for (int k = 0; k < kPW; ++k) {
                    if (k < jb) {
                        const int s = sm.src[k];
                        nw[k] = (s < jb) ? sm.T[s][tid] : col[s];
                    }
                }
#pragma unroll
                for (int k = 0; k < kPW; ++k)
                    if (k < jb && sm.dst[k] >= jb) col[sm.dst[k]] = sm.T[k][tid];
                if (c >= j0 + jb) {
#pragma unroll
                    for (int k = 0; k < kPW; ++k) {
                        const float u = -nw[k];
#pragma unroll
                        for (int i = k + 1; i < kPW; ++i) nw[i] = fmaf(sm.L11[i][k], u, nw[i]);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
#pragma unroll
                for (int k = 0; k < kPW; ++k)
                    if (k < jb) col[k] = nw[k];
            }
        }
        __syncthreads();
        // ---- trailing update on FP16 MFMA, split operands ----
        const int r0 = j0 + jb, c0 = j0 + jb;
        if (r0 < n) {
            const int wr = (w & 1) * 32, wc = (w >> 1) * 32;     // wave quadrant of the 64x64 tile
            const int li = lane & 15, lq = lane >> 4;
            for (int tc = c0; tc < n; tc += kTile) {
                for (int tr = r0; tr < n; tr += kTile) {
                    // stage L21 rows tr..tr+63 and U12 columns tc..tc+63 (k = jb), split
                    for (int e = tid; e < kTile * kPW; e += kT) {
                        const int i = e % kTile, k = e / kTile;
                        const int gi = tr + i;
                        const float v = (gi < n && k < jb) ? -W[gi + (size_t)(j0 + k) * n] : 0.0f;
                        const _Float16 h = (_Float16)v;
                        sm.Lh[i][k] = h;
                        sm.Ll[i][k] = (_Float16)(v - (float)h);
                    }
                    for (int e = tid; e < kTile * kPW; e += kT) {
                        const int k = e % kPW, c = e / kPW;
                        const int gc = tc + c;
                        const float v = (gc < n && k < jb) ? W[(j0 + k) + (size_t)gc * n] : 0.0f;
                        const _Float16 h = (_Float16)v;
                        sm.Uh[c][k] = h;
                        sm.Ul[c][k] = (_Float16)(v - (float)h);
                    }
                    __syncthreads();
                    float4v acc[2][2];
#pragma unroll
                    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                        for (int bj = 0; bj < 2; ++bj) {
                            const int col = tc + wc + 16 * bj + li;
#pragma unroll
                            for (int q = 0; q < 4; ++q) {
                                const int row = tr + wr + 16 * bi + 4 * lq + q;
                                acc[bi][bj][q] = (row < n && col < n) ? W[row + (size_t)col * n] : 0.0f;
                            }
                        }
#pragma unroll
                    for (int ks = 0; ks < kPW; ks += 16) {
                        half4 ah[2], al[2], bh[2], bl[2];
#pragma unroll
                        for (int bi = 0; bi < 2; ++bi) {
                            const int i = wr + 16 * bi + li;
                            ah[bi] = *reinterpret_cast<const half4 *>(&sm.Lh[i][ks + 4 * lq]);
                            al[bi] = *reinterpret_cast<const half4 *>(&sm.Ll[i][ks + 4 * lq]);
                        }
#pragma unroll
                        for (int bj = 0; bj < 2; ++bj) {
                            const int c = wc + 16 * bj + li;
                            bh[bj] = *reinterpret_cast<const half4 *>(&sm.Uh[c][ks + 4 * lq]);
                            bl[bj] = *reinterpret_cast<const half4 *>(&sm.Ul[c][ks + 4 * lq]);
                        }
#pragma unroll
                        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                            for (int bj = 0; bj < 2; ++bj) {
                                acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x16f16(al[bi], bh[bj], acc[bi][bj], 0, 0, 0);
                                acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah[bi], bl[bj], acc[bi][bj], 0, 0, 0);
                                acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah[bi], bh[bj], acc[bi][bj], 0, 0, 0);
                            }
                    }
#pragma unroll
                    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                        for (int bj = 0; bj < 2; ++bj) {
                            const int col = tc + wc + 16 * bj + li;
#pragma unroll
                            for (int q = 0; q < 4; ++q) {
                                const int row = tr + wr + 16 * bi + 4 * lq + q;
                                if (row < n && col < n) W[row + (size_t)col * n] = acc[bi][bj][q];
                            }
                        }
                    __syncthreads();
                }
            }
        }
    }
    return info;
}

// y (LDS, FP32, length n) := U^-1 L^-1 y, blocked by 32: a wavefront solves
// each diagonal block with shuffles, all threads apply the block to the rest.
__device__ __forceinline__ void trsv32(const float *W, int n, float *y)
{
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    // forward, unit lower
    for (int k0 = 0; k0 < n; k0 += 32) {
        const int kb = min(32, n - k0);
        if (tid < 64) {
            float v = (lane < kb) ? y[k0 + lane] : 0.0f;
            float row[32];
#pragma unroll
            for (int k = 0; k < 32; ++k) row[k] = (lane < kb && k < kb) ? W[(k0 + lane) + (size_t)(k0 + k) * n] : 0.0f;
#pragma unroll
            for (int k = 0; k < 32; ++k) {
                const float yk = __shfl(v, k);
                if (lane > k) v = fmaf(-row[k], yk, v);
            }
            if (lane < kb) y[k0 + lane] = v;
        }
        __syncthreads();
        for (int r = k0 + kb + tid; r < n; r += kT) {
            float v = y[r];
            for (int k = 0; k < kb; ++k) v = fmaf(-W[r + (size_t)(k0 + k) * n], y[k0 + k], v);
            y[r] = v;
        }
        __syncthreads();
    }
    // backward, upper
    for (int kend = n; kend > 0; kend -= 32) {
        const int k0 = max(0, kend - 32);
        const int kb = kend - k0;
        if (tid < 64) {
            float v = (lane < kb) ? y[k0 + lane] : 0.0f;
            float row[32];
#pragma unroll
            for (int k = 0; k < 32; ++k) row[k] = (lane < kb && k < kb) ? W[(k0 + lane) + (size_t)(k0 + k) * n] : 1.0f;
#pragma unroll
            for (int k = 31; k >= 0; --k) {
                if (k < kb) {
                    if (lane == k) v = v / row[k];
                    const float xk = __shfl(v, k);
                    if (lane < k) v = fmaf(-row[k], xk, v);
                }
            }
            if (lane < kb) y[k0 + lane] = v;
        }
        __syncthreads();
        for (int r = tid; r < k0; r += kT) {
            float v = y[r];
            for (int k = 0; k < kb; ++k) v = fmaf(-W[r + (size_t)(k0 + k) * n], y[k0 + k], v);
            y[r] = v;
        }
        __syncthreads();
    }
}

template <int RPT>
__global__ void __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(2))) dsgesv_kernel(MixedArgs g)
{
    __shared__ Smem sm;
    const int tid = threadIdx.x;
    const int n = g.n;
    float *W = g.W + (size_t)blockIdx.x * n * n;
    const double eps = 0x1p-53;
    for (int item = blockIdx.x; item < g.count; item += gridDim.x) {
        const double *A = g.A[item];
        const double *B = g.B[item];
        double *X = g.X[item];
        int *ipiv = g.ipiv[item];
        const int lda = g.lda;
        // ---- 1. norms and scaled FP32 copy ----
        double rowsum_max = 0.0, amax = 0.0;
        for (int r = tid; r < n; r += kT) {
            double s = 0.0, m = 0.0;
            for (int c = 0; c < n; ++c) {
                const double v = fabs(A[r + (size_t)c * lda]);
                s += v;
                m = nanmax(m, v);
            }
            rowsum_max = nanmax(rowsum_max, s);
            amax = nanmax(amax, m);
        }
        const double anrm = block_max(rowsum_max, sm.red);
        const double amx = block_max(amax, sm.red);
        int status = 0;           // 0 ok, -2 unscalable, -3 zero pivot, -(ITERMAX+1) no convergence
        if (!(amx > 0.0) || !__builtin_isfinite(amx)) {
            status = (amx == 0.0) ? -3 : -2;
        }
        int ex = 0;
        if (status == 0) {
            frexp(amx, &ex);
            const double sc = ldexp(1.0, -ex);
            for (int e = tid; e < n * n; e += kT) {
                const int r = e % n, c = e / n;
                W[e] = (float)(A[r + (size_t)c * lda] * sc);
            }
            __syncthreads();
            // ---- 2. FP32 LU ----
            if (lu32<RPT>(W, n, ipiv, sm) != 0) status = -3;
        }
        int iters = 0;
        if (status == 0) {
            // composed row permutation of the factorization: row i of P b is b[perm[i]]
            for (int i = tid; i < n; i += kT) sm.perm[i] = i;
            __syncthreads();
            if (tid == 0)
                for (int k = 0; k < n; ++k) {
                    const int p = ipiv[k] - 1;
                    const int t = sm.perm[k];
                    sm.perm[k] = sm.perm[p];
                    sm.perm[p] = t;
                }
            __syncthreads();
            const double cte = anrm * eps * sqrt((double)n);   // BWDMAX = 1
            const double inv = ldexp(1.0, -ex);                // A = 2^ex * (factored matrix)
            // columns are refined one after the other; iter is the largest
            // number of correction sweeps any column needed
            for (int j = 0; j < g.nrhs && status == 0; ++j) {
                const double *b = B + (size_t)j * g.ldb;
                double *x = X + (size_t)j * g.ldx;
                for (int i = tid; i < n; i += kT) {
                    sm.x[i] = 0.0;
                    sm.r[i] = b[i];
                }
                __syncthreads();
                bool conv = false;
                for (int it = 0; it <= MAGMA_MIXED_ITERMAX; ++it) {
                    // d = LU32^-1 (P r) in FP32, x += d in FP64
                    for (int i = tid; i < n; i += kT) sm.y[i] = (float)(sm.r[sm.perm[i]] * inv);
                    __syncthreads();
                    trsv32(W, n, sm.y);
                    for (int i = tid; i < n; i += kT) sm.x[i] += (double)sm.y[i];
                    __syncthreads();
                    // r = b - A x in FP64
                    double rmax = 0.0, xmax = 0.0;
                    for (int r = tid; r < n; r += kT) {
                        double acc = b[r];
                        for (int c = 0; c < n; ++c) acc = fma(-A[r + (size_t)c * lda], sm.x[c], acc);
                        sm.r[r] = acc;
                        const double ar = fabs(acc), ax = fabs(sm.x[r]);
                        rmax = (ar > rmax || ar != ar) ? ar : rmax;
                        xmax = (ax > xmax || ax != ax) ? ax : xmax;
                    }
                    const double rn = block_max(rmax, sm.red);
                    const double xn = block_max(xmax, sm.red);
                    if (rn <= xn * cte) {
                        conv = true;
                        iters = max(iters, it);
                        break;
                    }
                }
                if (!conv) status = -(MAGMA_MIXED_ITERMAX + 1);
                for (int i = tid; i < n; i += kT) x[i] = sm.x[i];
                __syncthreads();
            }
        }
        if (status != 0 && tid == 0) {
            const int slot = atomicAdd(g.fb_count, 1);
            g.fb_list[slot] = item;
        }
        if (tid == 0) {
            g.iter[item] = status == 0 ? iters : status;
            g.info[item] = 0;
        }
        __syncthreads();
    }
}

// X <- B for the matrices that fall back (the FP64 solve overwrites X in place)
__global__ void copy_rhs_kernel(const int *list, const int *count, double **B, int ldb, double **X, int ldx,
                                int n, int nrhs)
{
    const int cnt = *count;
    for (int t = blockIdx.x; t < cnt; t += gridDim.x) {
        const int id = list[t];
        for (int e = threadIdx.x; e < n * nrhs; e += blockDim.x) {
            const int i = e % n, c = e / n;
            X[id][i + (size_t)c * ldx] = B[id][i + (size_t)c * ldb];
        }
    }
}

__global__ void zero_one_kernel(int *p) { *p = 0; }

__global__ void fill_fallback_all_kernel(int *list, int *count, int *iter, int batch)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < batch) {
        list[i] = i;
        iter[i] = -1;
    }
    if (i == 0) *count = batch;
}

}  // namespace
}  // namespace magma_amd

using namespace magma_amd;

extern "C" magma_int_t magma_dsgesv_batched(magma_int_t n, magma_int_t nrhs,
                                            double **dA_array, magma_int_t ldda, magma_int_t **dipiv_array,
                                            double **dB_array, magma_int_t lddb,
                                            double **dX_array, magma_int_t lddx,
                                            magma_int_t *dinfo_array, magma_int_t *diter_array,
                                            magma_int_t batchCount, magma_queue_t queue)
{
    // argument checks in the order of magma_dsgesv_gpu (src/zcgesv_gpu.cpp:163-180)
    magma_int_t info = 0;
    if (n < 0) info = -1;
    else if (nrhs < 0) info = -2;
    else if (ldda < std::max(1, n)) info = -4;
    else if (lddb < std::max(1, n)) info = -7;
    else if (lddx < std::max(1, n)) info = -9;
    else if (batchCount < 0) info = -12;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    if (n == 0 || nrhs == 0 || batchCount == 0) return MAGMA_SUCCESS;
    hipStream_t st = stream_of(queue);
    const int grid = persistent_grid(batchCount, 4);
    const bool mixed_ok = n <= kMaxN;
    const size_t wbytes = mixed_ok ? sizeof(float) * (size_t)n * n * grid : 0;
    const size_t lbytes = sizeof(int) * (size_t)(batchCount + 1);
    char *ws = static_cast<char *>(queue_workspace(queue, round_up(wbytes, 256) + lbytes));
    if (!ws) return MAGMA_ERR_DEVICE_ALLOC;
    int *fb_count = reinterpret_cast<int *>(ws + round_up(wbytes, 256));
    int *fb_list = fb_count + 1;
    zero_one_kernel<<<1, 1, 0, st>>>(fb_count);
    MixedArgs g;
    g.n = n; g.nrhs = nrhs; g.lda = ldda; g.ldb = lddb; g.ldx = lddx;
    g.A = dA_array; g.B = dB_array; g.X = dX_array;
    g.ipiv = dipiv_array; g.info = dinfo_array; g.iter = diter_array;
    g.W = reinterpret_cast<float *>(ws);
    g.fb_list = fb_list; g.fb_count = fb_count;
    g.count = batchCount;
    if (mixed_ok) {
        if (n <= kT) dsgesv_kernel<1><<<grid, kT, 0, st>>>(g);
        else dsgesv_kernel<2><<<grid, kT, 0, st>>>(g);
        MAGMA_LAUNCH_CHECK();
    } else {
        // n above the mixed kernel's range: every matrix takes the FP64 path
        // (dsgesv's iter = -1, "implementation-specific reasons")
        fill_fallback_all_kernel<<<ceil_div(batchCount, 256), 256, 0, st>>>(fb_list, fb_count, diter_array,
                                                                             batchCount);
        MAGMA_LAUNCH_CHECK();
    }
    // FP64 fallback on the device list: X <- B, getrf(A), getrs(X)
    copy_rhs_kernel<<<persistent_grid(batchCount, 4), 256, 0, st>>>(fb_list, fb_count, dB_array, lddb, dX_array,
                                                                    lddx, n, nrhs);
    MatSet s;
    s.m = n; s.n = n; s.lda = ldda;
    s.A = dA_array; s.ipiv = dipiv_array; s.info = dinfo_array;
    s.B = dX_array; s.nrhs = nrhs; s.ldb = lddx;
    s.index = fb_list; s.dev_count = fb_count; s.count = batchCount;
    magma_int_t r = launch_getrf(s, n, n, true, st);
    if (r != MAGMA_SUCCESS) return r;
    if (!queue->async) MAGMA_HIP_CHECK(hipStreamSynchronize(st));
    return MAGMA_SUCCESS;
}
