// gemm.hip -- batched and vbatched DGEMM, C = alpha op(A) op(B) + beta C, on
// FP64 MFMA (v_mfma_f64_16x16x4f64).
//
// Each wavefront owns one 16x16, 32x32 or 48x48 tile of C (1, 4 or 9 MFMA
// blocks of 16x16; the variant is tuned per size class) and
// walks k four at a time; a workgroup holds 4 independent wavefronts.  A
// fourth variant gives a workgroup one 64x64 tile with its operands staged
// through LDS (gemm_lds_kernel).  The
// grid is persistent: wavefronts stride over a flat tile list.  For vbatched
// calls the list is the device prefix sum of per-matrix tile counts
// (workqueue.hip), so a matrix contributes exactly its own tiles instead of a
// grid sized by the largest matrix with most blocks exiting
// (magmablas/gemm_template_kernel_vbatched.cuh:18-304,
// magmablas/dgemm_vbatched_core.cu:27-266).
//
// MFMA operand layout (v_mfma_f64_16x16x4f64, wave64): lane l supplies
// A(l%16, l/16) and B(l/16, l%16) of the 16x4 / 4x16 blocks and receives
// D(l/16 + 4*i, l%16), i = 0..3, of the 16x16 result (FP64 interleaves the
// result rows, unlike the FP16/FP32 variants' 4*(l/16) + i).
#include <type_traits>

#include "kernels.h"
#include "workqueue.h"

namespace magma_amd {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;

__device__ __forceinline__ int pick(const int *arr, int fixed, int id) { return arr ? arr[id] : fixed; }

// p[id] read as a global (address space 1) pointer and converted back, so
// the compiler emits global rather than flat loads and stores through it.  A
// flat load also counts against the LDS wait counter: the next k block's
// operand loads, issued ahead of the current block's MFMAs, would be waited
// for at the first LDS read of the current block.
template <class T>
__device__ __forceinline__ T *global_ptr(T *const *p, int id)
{
    typedef typename std::remove_const<T>::type U;
    typedef U __attribute__((address_space(1))) *G;
    return (T *)((G const *)p)[id];
}

template <bool TA, bool TB, int BT>
__global__ void __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(BT == 1 ? 4 : (BT == 2 ? 3 : 2)))) gemm_mfma_kernel(GemmSet g, const int *prefix, int fixed_tiles)
{
    constexpr int TS = 16 * BT;        // tile edge: BT x BT MFMA blocks of 16x16
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int li = lane & 15;          // row within a 16-block (A) / column (B, D)
    const int lk = lane >> 4;          // k within the 4-slice (A, B) / row quad (D)
    const int total = prefix ? prefix[g.count] : fixed_tiles * g.count;
    const int stride = gridDim.x * kWaves;

    for (int t = blockIdx.x * kWaves + wave; t < total; t += stride) {
        int s, local;
        if (prefix) {
            int lo = 0, hi = g.count;          // largest s with prefix[s] <= t
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (prefix[mid] <= t) lo = mid; else hi = mid;
            }
            s = lo;
            local = t - prefix[s];
        } else {
            s = t / fixed_tiles;
            local = t - s * fixed_tiles;
        }
        const int id = g.index ? g.index[s] : s;
        const int m = pick(g.m_arr, g.m, id), n = pick(g.n_arr, g.n, id), k = pick(g.k_arr, g.k, id);
        const int lda = pick(g.lda_arr, g.lda, id), ldb = pick(g.ldb_arr, g.ldb, id);
        const int ldc = pick(g.ldc_arr, g.ldc, id);
        const int mt = (m + TS - 1) / TS;
        const int row0 = (local % mt) * TS;
        const int col0 = (local / mt) * TS;
        const double *A = global_ptr(g.A, id);
        const double *B = global_ptr(g.B, id);
        double *C = global_ptr(g.C, id);

        d4 acc[BT][BT];
#pragma unroll
        for (int bi = 0; bi < BT; ++bi)
#pragma unroll
            for (int bj = 0; bj < BT; ++bj) acc[bi][bj] = d4{0.0, 0.0, 0.0, 0.0};

        // Main loop, 16 k per step, the next step's operands loaded while the
        // current step runs: in slice s lane (li, lk) supplies k = k0 + 4 lk + s
        // (a permutation of k inside the step), so a lane's four k of B (and of
        // A^T) are one contiguous 32-byte piece.  Rows and columns past the edge
        // load a valid address and multiply by zero.
        const int kmain = k & ~15;
        int rq[BT], cq[BT];
#pragma unroll
        for (int q = 0; q < BT; ++q) {
            rq[q] = min(row0 + 16 * q + li, m - 1);
            cq[q] = min(col0 + 16 * q + li, n - 1);
        }
        double av[2][BT][4], bv[2][BT][4];
        auto load16 = [&](int kb, double (&ao)[BT][4], double (&bo)[BT][4]) {
#pragma unroll
            for (int q = 0; q < BT; ++q) {
                const int r = rq[q], c = cq[q];
                const bool rin = row0 + 16 * q + li < m, cin = col0 + 16 * q + li < n;
#pragma unroll
                for (int sl = 0; sl < 4; ++sl) {
                    const int kk = kb + 4 * lk + sl;
                    const double av_ = TA ? A[kk + (size_t)r * lda] : A[r + (size_t)kk * lda];
                    const double bv_ = TB ? B[c + (size_t)kk * ldb] : B[kk + (size_t)c * ldb];
                    ao[q][sl] = rin ? av_ : 0.0;
                    bo[q][sl] = cin ? bv_ : 0.0;
                }
            }
        };
        if (kmain > 0) load16(0, av[0], bv[0]);
#pragma unroll 1
        for (int k0 = 0; k0 < kmain; k0 += 32) {
            if (k0 + 16 < kmain) load16(k0 + 16, av[1], bv[1]);
#pragma unroll
            for (int sl = 0; sl < 4; ++sl)
#pragma unroll
                for (int bi = 0; bi < BT; ++bi)
#pragma unroll
                    for (int bj = 0; bj < BT; ++bj)
                        acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[0][bi][sl], bv[0][bj][sl], acc[bi][bj], 0, 0, 0);
            if (k0 + 16 >= kmain) break;
            if (k0 + 32 < kmain) load16(k0 + 32, av[0], bv[0]);
#pragma unroll
            for (int sl = 0; sl < 4; ++sl)
#pragma unroll
                for (int bi = 0; bi < BT; ++bi)
#pragma unroll
                    for (int bj = 0; bj < BT; ++bj)
                        acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[1][bi][sl], bv[1][bj][sl], acc[bi][bj], 0, 0, 0);
        }
        // tail, 4 k per step in order
        for (int k0 = kmain; k0 < k; k0 += 4) {
            const int kk = k0 + lk;
            const bool kin = kk < k;
            double a[BT], b[BT];
#pragma unroll
            for (int q = 0; q < BT; ++q) {
                const int r = row0 + 16 * q + li, c = col0 + 16 * q + li;
                a[q] = 0.0;
                b[q] = 0.0;
                if (kin) {
                    if (r < m) a[q] = TA ? A[kk + (size_t)r * lda] : A[r + (size_t)kk * lda];
                    if (c < n) b[q] = TB ? B[c + (size_t)kk * ldb] : B[kk + (size_t)c * ldb];
                }
            }
#pragma unroll
            for (int bi = 0; bi < BT; ++bi)
#pragma unroll
                for (int bj = 0; bj < BT; ++bj)
                    acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], b[bj], acc[bi][bj], 0, 0, 0);
        }
        const double alpha = g.alpha, beta = g.beta;
#pragma unroll
        for (int bi = 0; bi < BT; ++bi) {
#pragma unroll
            for (int bj = 0; bj < BT; ++bj) {
                const int col = col0 + 16 * bj + li;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = row0 + 16 * bi + lk + 4 * i;
                    if (row < m && col < n) {
                        double *c = C + row + (size_t)col * ldc;
                        // beta == 0: C is not read (BLAS semantics)
                        *c = (beta == 0.0) ? alpha * acc[bi][bj][i] : fma(alpha, acc[bi][bj][i], beta * *c);
                    }
                }
            }
        }
    }
}

// Variant with a 64 x 64 C tile per workgroup.  The k blocks of A and B pass
// through LDS (double-buffered, one barrier per 16 k), so a workgroup reads
// each operand once instead of once per wavefront; the four wavefronts own
// 32 x 32 quadrants.  Every entry is accumulated in the same k order as the
// per-wavefront tiles (16-k steps with the in-step permutation
// k0 + 4 lk + sl, then the tail four at a time), so the variants agree bit for
// bit.
constexpr int kLT = 64;           // workgroup tile edge
constexpr int kLK = 16;           // k per stage
constexpr int kLPitch = kLT + 4;  // LDS row pitch, in doubles

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
gemm_lds_kernel(GemmSet g, const int *prefix, int fixed_tiles)
{
    __shared__ double As[2][kLK][kLPitch];   // [stage][k][row within the tile]
    __shared__ double Bs[2][kLK][kLPitch];   // [stage][k][column within the tile]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int wr = wave & 1, wc = wave >> 1;  // this wavefront's quadrant
    // element of a stage each thread moves: A (row ar + 16 j or + 4 j, k ak),
    // B (k bk, column bc), chosen so the global loads are unit-stride
    const int ar = TA ? tid / kLK : tid % kLT, ak = TA ? tid % kLK : tid / kLT;
    const int bc = TB ? tid % kLT : tid / kLK, bk = TB ? tid / kLT : tid % kLK;
    constexpr int AST = TA ? 256 / kLK : 0, AKT = TA ? 0 : 256 / kLT;   // strides over j
    constexpr int BST = TB ? 0 : 256 / kLK, BKT = TB ? 256 / kLT : 0;
    const int total = prefix ? prefix[g.count] : fixed_tiles * g.count;

    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        int s, local;
        if (prefix) {
            int lo = 0, hi = g.count;          // largest s with prefix[s] <= t
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (prefix[mid] <= t) lo = mid; else hi = mid;
            }
            s = lo;
            local = t - prefix[s];
        } else {
            s = t / fixed_tiles;
            local = t - s * fixed_tiles;
        }
        const int id = g.index ? g.index[s] : s;
        const int m = pick(g.m_arr, g.m, id), n = pick(g.n_arr, g.n, id), k = pick(g.k_arr, g.k, id);
        const int lda = pick(g.lda_arr, g.lda, id), ldb = pick(g.ldb_arr, g.ldb, id);
        const int ldc = pick(g.ldc_arr, g.ldc, id);
        const int mt = (m + kLT - 1) / kLT;
        const int row0 = (local % mt) * kLT;
        const int col0 = (local / mt) * kLT;
        const double *A = global_ptr(g.A, id);
        const double *B = global_ptr(g.B, id);
        double *C = global_ptr(g.C, id);

        d4 acc[2][2];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = d4{0.0, 0.0, 0.0, 0.0};

        const int kmain = k & ~(kLK - 1);
        double ra[4], rb[4];
        auto gload = [&](int kb) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int r = row0 + ar + AST * j, kk = kb + ak + AKT * j;
                const int c = col0 + bc + BST * j, kq = kb + bk + BKT * j;
                ra[j] = (r < m) ? (TA ? A[kk + (size_t)r * lda] : A[r + (size_t)kk * lda]) : 0.0;
                rb[j] = (c < n) ? (TB ? B[c + (size_t)kq * ldb] : B[kq + (size_t)c * ldb]) : 0.0;
            }
        };
        auto sstore = [&](int st) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                As[st][ak + AKT * j][ar + AST * j] = ra[j];
                Bs[st][bk + BKT * j][bc + BST * j] = rb[j];
            }
        };
        if (kmain > 0) {
            gload(0);
            sstore(0);
        }
        __syncthreads();
        int st = 0;
#pragma unroll 1
        for (int kb = 0; kb < kmain; kb += kLK) {
            const bool more = kb + kLK < kmain;
            if (more) gload(kb + kLK);
#pragma unroll
            for (int sl = 0; sl < 4; ++sl) {
                const int kk = 4 * lk + sl;
                double a[2], b[2];
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    a[q] = As[st][kk][32 * wr + 16 * q + li];
                    b[q] = Bs[st][kk][32 * wc + 16 * q + li];
                }
#pragma unroll
                for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 2; ++bj)
                        acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], b[bj], acc[bi][bj], 0, 0, 0);
            }
            if (more) sstore(st ^ 1);
            __syncthreads();
            st ^= 1;
        }
        // tail, 4 k per step in order (as the per-wavefront tiles)
        for (int k0 = kmain; k0 < k; k0 += 4) {
            const int kk = k0 + lk;
            const bool kin = kk < k;
            double a[2], b[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int r = row0 + 32 * wr + 16 * q + li, c = col0 + 32 * wc + 16 * q + li;
                a[q] = 0.0;
                b[q] = 0.0;
                if (kin) {
                    if (r < m) a[q] = TA ? A[kk + (size_t)r * lda] : A[r + (size_t)kk * lda];
                    if (c < n) b[q] = TB ? B[c + (size_t)kk * ldb] : B[kk + (size_t)c * ldb];
                }
            }
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
                    acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], b[bj], acc[bi][bj], 0, 0, 0);
        }
        const double alpha = g.alpha, beta = g.beta;
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                const int col = col0 + 32 * wc + 16 * bj + li;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = row0 + 32 * wr + 16 * bi + lk + 4 * i;
                    if (row < m && col < n) {
                        double *c = C + row + (size_t)col * ldc;
                        // beta == 0: C is not read (BLAS semantics)
                        *c = (beta == 0.0) ? alpha * acc[bi][bj][i] : fma(alpha, acc[bi][bj][i], beta * *c);
                    }
                }
            }
        }
        // the next tile's first stage overwrites stage 0
        __syncthreads();
    }
}

// C = beta C when k == 0 or alpha == 0 (BLAS quick path)
__global__ void gemm_scale_kernel(GemmSet g, const int *prefix, int fixed_tiles, int ts)
{
    const int total = prefix ? prefix[g.count] : fixed_tiles * g.count;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        int s, local;
        if (prefix) {
            int lo = 0, hi = g.count;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (prefix[mid] <= t) lo = mid; else hi = mid;
            }
            s = lo;
            local = t - prefix[s];
        } else {
            s = t / fixed_tiles;
            local = t - s * fixed_tiles;
        }
        const int id = g.index ? g.index[s] : s;
        const int m = pick(g.m_arr, g.m, id), n = pick(g.n_arr, g.n, id);
        const int ldc = pick(g.ldc_arr, g.ldc, id);
        const int mt = (m + ts - 1) / ts;
        const int row0 = (local % mt) * ts, col0 = (local / mt) * ts;
        double *C = g.C[id];
        for (int e = threadIdx.x; e < ts * ts; e += blockDim.x) {
            const int row = row0 + e % ts, col = col0 + e / ts;
            if (row < m && col < n) {
                double *c = C + row + (size_t)col * ldc;
                *c = (g.beta == 0.0) ? 0.0 : g.beta * *c;
            }
        }
    }
}

}  // namespace

template <int BT>
static void launch_bt(const GemmSet &g, const int *prefix, int fixed_tiles, int grid, hipStream_t st)
{
    if (g.ta == 'N' && g.tb == 'N') gemm_mfma_kernel<false, false, BT><<<grid, 64 * kWaves, 0, st>>>(g, prefix, fixed_tiles);
    else if (g.ta == 'N') gemm_mfma_kernel<false, true, BT><<<grid, 64 * kWaves, 0, st>>>(g, prefix, fixed_tiles);
    else if (g.tb == 'N') gemm_mfma_kernel<true, false, BT><<<grid, 64 * kWaves, 0, st>>>(g, prefix, fixed_tiles);
    else gemm_mfma_kernel<true, true, BT><<<grid, 64 * kWaves, 0, st>>>(g, prefix, fixed_tiles);
}

// prefix: device tile prefix (vbatched) or null (fixed sizes), built for tile
// edge `tile` (16, 32, 48: one tile per wavefront; 64: one per workgroup).
// max_tiles bounds the flat tile count for the grid size.
magma_int_t launch_gemm_tiles(const GemmSet &g, const int *prefix, long max_tiles, int tile, hipStream_t st)
{
    if (g.count <= 0 || max_tiles <= 0) return MAGMA_SUCCESS;
    if (tile != 16 && tile != 32 && tile != 48 && tile != kLT) return MAGMA_ERR_NOT_SUPPORTED;
    const int fixed_tiles = prefix ? 0 : ceil_div(g.m, tile) * ceil_div(g.n, tile);
    // enough resident wavefronts to fill 256 CUs several times over
    const long want = tile == kLT ? max_tiles : (max_tiles + kWaves - 1) / kWaves;
    const int grid = (int)std::min<long>(want, 256L * 16);
    const bool scale_only = (!prefix && (g.k == 0 || g.alpha == 0.0)) || g.alpha == 0.0;
    if (scale_only) {
        gemm_scale_kernel<<<grid, 256, 0, st>>>(g, prefix, fixed_tiles, tile);
    } else if (tile == kLT) {
        if (g.ta == 'N' && g.tb == 'N') gemm_lds_kernel<false, false><<<grid, 256, 0, st>>>(g, prefix, fixed_tiles);
        else if (g.ta == 'N') gemm_lds_kernel<false, true><<<grid, 256, 0, st>>>(g, prefix, fixed_tiles);
        else if (g.tb == 'N') gemm_lds_kernel<true, false><<<grid, 256, 0, st>>>(g, prefix, fixed_tiles);
        else gemm_lds_kernel<true, true><<<grid, 256, 0, st>>>(g, prefix, fixed_tiles);
    } else if (tile == 48) {
        launch_bt<3>(g, prefix, fixed_tiles, grid, st);
    } else if (tile == 32) {
        launch_bt<2>(g, prefix, fixed_tiles, grid, st);
    } else {
        launch_bt<1>(g, prefix, fixed_tiles, grid, st);
    }
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

}  // namespace magma_amd
