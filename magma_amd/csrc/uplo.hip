// uplo.hip -- MagmaUpper Cholesky through the lower-triangular kernels.
//
// For a symmetric matrix stored in its upper triangle, the element (i, j),
// i > j, of the lower triangle of A' is A(j, i).  Swapping every pair
// A(i, j) <-> A(j, i) in place therefore hands the lower-triangular kernels
// exactly the values they would read from a lower-stored copy; they factor
// A' = L L' with the netlib operation order, and the second swap leaves
// U = L' in the upper triangle and the caller's strictly lower triangle as
// it was.  Element for element this is the arithmetic of LAPACK's upper
// dpotrf (U(j, i) = (a_ji - sum_k u_kj u_ki) / u_jj, the mirror image of the
// lower case), so info and the factor are those of dpotrf('U').  The
// reference rejects Upper in its batched Cholesky
// (src/zpotrf_batched.cpp:106-109, magmablas/zpotf2_kernels.cu:153-156).
#include "kernels.h"

namespace magma_amd {

namespace {

constexpr int kTile = 32;
constexpr int kSwapThreads = 256;

__device__ __forceinline__ int pick(const int *arr, int fixed, int id) { return arr ? arr[id] : fixed; }

// One workgroup per matrix (persistent).  The tile pairs (ti, tj), ti >= tj,
// are staged through LDS so both the loads and the stores run down columns.
__global__ void __launch_bounds__(kSwapThreads) swap_triangles_kernel(MatSet s)
{
    __shared__ double t1[kTile][kTile + 1];
    __shared__ double t2[kTile][kTile + 1];
    const int cnt = item_count(s);
    const int *idx = item_index(s);
    const int r = threadIdx.x % kTile, cg = threadIdx.x / kTile;     // row in tile, column group
    constexpr int kCols = kTile / (kSwapThreads / kTile);            // columns per thread
    for (int item = blockIdx.x; item < cnt; item += gridDim.x) {
        const int id = idx ? idx[item] : item;
        const int n = pick(s.n_arr, s.n, id);
        const int lda = pick(s.lda_arr, s.lda, id);
        if (n <= 1 || lda < n) continue;
        double *A = s.A[id];
        const int T = (n + kTile - 1) / kTile;
        for (int ti = 0; ti < T; ++ti) {
            for (int tj = 0; tj <= ti; ++tj) {
                const int i0 = ti * kTile, j0 = tj * kTile;
                // t1 = A(i0.., j0..) (lower tile), t2 = A(j0.., i0..) (upper tile)
#pragma unroll
                for (int q = 0; q < kCols; ++q) {
                    const int c = cg + (kSwapThreads / kTile) * q;
                    const bool in1 = i0 + r < n && j0 + c < n;
                    const bool in2 = j0 + r < n && i0 + c < n;
                    t1[c][r] = in1 ? A[(i0 + r) + (size_t)(j0 + c) * lda] : 0.0;
                    t2[c][r] = (in2 && ti != tj) ? A[(j0 + r) + (size_t)(i0 + c) * lda] : 0.0;
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < kCols; ++q) {
                    const int c = cg + (kSwapThreads / kTile) * q;
                    if (ti == tj) {
                        // diagonal tile: its transpose (the diagonal stays)
                        if (i0 + r < n && j0 + c < n && r != c) A[(i0 + r) + (size_t)(j0 + c) * lda] = t1[r][c];
                    } else {
                        if (i0 + r < n && j0 + c < n) A[(i0 + r) + (size_t)(j0 + c) * lda] = t2[r][c];
                        if (j0 + r < n && i0 + c < n) A[(j0 + r) + (size_t)(i0 + c) * lda] = t1[r][c];
                    }
                }
                __syncthreads();
            }
        }
    }
}

}  // namespace

magma_int_t launch_swap_triangles(const MatSet &s, hipStream_t st)
{
    if (s.count <= 0) return MAGMA_SUCCESS;
    swap_triangles_kernel<<<persistent_grid(s.count, 8), kSwapThreads, 0, st>>>(s);
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

}  // namespace magma_amd
