// drivers.hip -- the magma_d*_batched C entry points: argument checking with
// the reference's error positions, size-class dispatch to the kernel families,
// and the reference's synchronous completion.
//
// Dispatch (replaces the static crossover tables of
// control/get_batched_crossover.cpp:25-228 and control/batched_kernel_param.h):
//   max(m,n) <= 64   register-resident kernels, 64/NMAX matrices per wavefront
//   max(m,n) <= 512  LU: workgroup per matrix, FP64 MFMA trailing update
//   otherwise        blocked workgroup-per-matrix kernels (VALU)
#include <algorithm>

#include "kernels.h"

namespace magma_amd {

magma_int_t launch_getrf_small(const MatSet &s, int max_mn_dim, bool solve, hipStream_t st);
magma_int_t launch_potrf_small(const MatSet &s, int max_n, bool solve, hipStream_t st);
magma_int_t launch_getrf_blocked(const MatSet &s, int max_m, hipStream_t st);
magma_int_t launch_getrf_mfma(const MatSet &s, int max_dim, hipStream_t st);
magma_int_t launch_potrf_blocked(const MatSet &s, int max_n, hipStream_t st);

constexpr int kRegMax = 64;
constexpr int kMfmaMax = 512;

namespace {
__global__ void fill_int_kernel(int *p, int v, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}
}  // namespace

// Sets info_array[0..count) to zero with a kernel (safe under stream capture).
magma_int_t zero_info(int *info, int count, hipStream_t st)
{
    if (!info || count <= 0) return MAGMA_SUCCESS;
    fill_int_kernel<<<ceil_div(count, 256), 256, 0, st>>>(info, 0, count);
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

magma_int_t launch_getrf(const MatSet &s, int max_m, int max_n, bool fused_solve, hipStream_t st)
{
    const int dim = std::max(max_m, max_n);
    if (dim <= kRegMax) return launch_getrf_small(s, dim, fused_solve, st);
    magma_int_t r = (dim <= kMfmaMax) ? launch_getrf_mfma(s, dim, st) : launch_getrf_blocked(s, max_m, st);
    if (r != MAGMA_SUCCESS || !fused_solve) return r;
    return launch_getrs(s, 'N', max_n, 0, st);
}

magma_int_t launch_potrf(const MatSet &s, int max_n, bool fused_solve, hipStream_t st)
{
    if (max_n <= kRegMax) return launch_potrf_small(s, max_n, fused_solve, st);
    magma_int_t r = launch_potrf_blocked(s, max_n, st);
    if (r != MAGMA_SUCCESS || !fused_solve) return r;
    return launch_potrs(s, 'L', max_n, 0, st);
}

static char trans_char(magma_trans_t t) { return t == MagmaNoTrans ? 'N' : 'T'; }

static magma_int_t finish(magma_int_t r, magma_queue_t queue)
{
    if (r == MAGMA_SUCCESS && !(queue && queue->async)) {
        hipError_t e = hipStreamSynchronize(stream_of(queue));
        if (e != hipSuccess) return report_hip_error(e, "hipStreamSynchronize", __FILE__, __LINE__);
    }
    return r;
}

}  // namespace magma_amd

using namespace magma_amd;

extern "C" {

// src/zgetrf_batched.cpp:82-121 (argument checks and quick return)
magma_int_t magma_dgetrf_batched(magma_int_t m, magma_int_t n, double **dA_array, magma_int_t ldda,
                                 magma_int_t **ipiv_array, magma_int_t *info_array,
                                 magma_int_t batchCount, magma_queue_t queue)
{
    // info_array is zeroed on entry as the reference does
    // (src/zgetrf_batched.cpp:94); on the factorization path the kernels
    // write every entry, so only the early returns clear it.
    hipStream_t st = stream_of(queue);
    magma_int_t arginfo = 0;
    if (m < 0) arginfo = -1;
    else if (n < 0) arginfo = -2;
    else if (ldda < std::max(1, m)) arginfo = -4;
    if (arginfo != 0) {
        zero_info(info_array, batchCount, st);
        magma_xerbla(__func__, -arginfo);
        return arginfo;
    }
    if (m == 0 || n == 0 || batchCount <= 0) {
        zero_info(info_array, batchCount, st);
        return finish(MAGMA_SUCCESS, queue);
    }
    MatSet s;
    s.m = m; s.n = n; s.lda = ldda;
    s.A = dA_array; s.ipiv = ipiv_array; s.info = info_array;
    s.count = batchCount;
    return finish(launch_getrf(s, m, n, false, st), queue);
}

// src/zgetrs_batched.cpp:86-118
magma_int_t magma_dgetrs_batched(magma_trans_t trans, magma_int_t n, magma_int_t nrhs,
                                 double **dA_array, magma_int_t ldda, magma_int_t **dipiv_array,
                                 double **dB_array, magma_int_t lddb,
                                 magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t info = 0;
    if (trans != MagmaNoTrans && trans != MagmaTrans && trans != MagmaConjTrans) info = -1;
    else if (n < 0) info = -2;
    else if (nrhs < 0) info = -3;
    else if (ldda < std::max(1, n)) info = -5;
    else if (lddb < std::max(1, n)) info = -8;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    if (n == 0 || nrhs == 0 || batchCount <= 0) return MAGMA_SUCCESS;
    MatSet s;
    s.m = n; s.n = n; s.lda = ldda;
    s.A = dA_array; s.ipiv = dipiv_array;
    s.B = dB_array; s.nrhs = nrhs; s.ldb = lddb;
    s.count = batchCount;
    return finish(launch_getrs(s, trans_char(trans), n, nrhs, stream_of(queue)), queue);
}

// src/zgesv_batched.cpp:92-147
magma_int_t magma_dgesv_batched(magma_int_t n, magma_int_t nrhs, double **dA_array, magma_int_t ldda,
                                magma_int_t **dipiv_array, double **dB_array, magma_int_t lddb,
                                magma_int_t *dinfo_array, magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t info = 0;
    if (n < 0) info = -1;
    else if (nrhs < 0) info = -2;
    else if (ldda < std::max(1, n)) info = -4;
    else if (lddb < std::max(1, n)) info = -6;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    hipStream_t st = stream_of(queue);
    if (n == 0 || nrhs == 0 || batchCount <= 0) {
        zero_info(dinfo_array, batchCount, st);
        return finish(MAGMA_SUCCESS, queue);
    }
    MatSet s;
    s.m = n; s.n = n; s.lda = ldda;
    s.A = dA_array; s.ipiv = dipiv_array; s.info = dinfo_array;
    s.B = dB_array; s.nrhs = nrhs; s.ldb = lddb;
    s.count = batchCount;
    return finish(launch_getrf(s, n, n, true, st), queue);
}

// src/zpotrf_batched.cpp:290-341.  The reference rejects Upper
// (src/zpotrf_batched.cpp:106-109); here it is LAPACK's upper dpotrf, through
// the lower kernels between two in-place triangle swaps (uplo.hip).
magma_int_t magma_dpotrf_batched(magma_uplo_t uplo, magma_int_t n, double **dA_array, magma_int_t lda,
                                 magma_int_t *info_array, magma_int_t batchCount, magma_queue_t queue)
{
    hipStream_t st = stream_of(queue);
    magma_int_t arginfo = 0;
    if (uplo != MagmaUpper && uplo != MagmaLower) arginfo = -1;
    else if (n < 0) arginfo = -2;
    else if (lda < std::max(1, n)) arginfo = -4;
    if (arginfo != 0) {
        zero_info(info_array, batchCount, st);
        magma_xerbla(__func__, -arginfo);
        return arginfo;
    }
    if (n == 0 || batchCount <= 0) {
        zero_info(info_array, batchCount, st);
        return finish(MAGMA_SUCCESS, queue);
    }
    MatSet s;
    s.n = n; s.m = n; s.lda = lda;
    s.A = dA_array; s.info = info_array;
    s.count = batchCount;
    if (uplo == MagmaUpper) {
        magma_int_t r = launch_swap_triangles(s, st);
        if (r == MAGMA_SUCCESS) r = launch_potrf(s, n, false, st);
        if (r == MAGMA_SUCCESS) r = launch_swap_triangles(s, st);
        return finish(r, queue);
    }
    return finish(launch_potrf(s, n, false, st), queue);
}

// src/zpotrs_batched.cpp:70-103
magma_int_t magma_dpotrs_batched(magma_uplo_t uplo, magma_int_t n, magma_int_t nrhs,
                                 double **dA_array, magma_int_t ldda, double **dB_array, magma_int_t lddb,
                                 magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t info = 0;
    if (uplo != MagmaUpper && uplo != MagmaLower) info = -1;
    if (n < 0) info = -2;
    if (nrhs < 0) info = -3;
    if (ldda < std::max(1, n)) info = -5;
    if (lddb < std::max(1, n)) info = -7;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    if (n == 0 || nrhs == 0 || batchCount <= 0) return MAGMA_SUCCESS;
    MatSet s;
    s.m = n; s.n = n; s.lda = ldda;
    s.A = dA_array;
    s.B = dB_array; s.nrhs = nrhs; s.ldb = lddb;
    s.count = batchCount;
    return finish(launch_potrs(s, uplo == MagmaLower ? 'L' : 'U', n, nrhs, stream_of(queue)), queue);
}

// src/zposv_batched.cpp:88-145
magma_int_t magma_dposv_batched(magma_uplo_t uplo, magma_int_t n, magma_int_t nrhs,
                                double **dA_array, magma_int_t ldda, double **dB_array, magma_int_t lddb,
                                magma_int_t *dinfo_array, magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t info = 0;
    if (uplo != MagmaUpper && uplo != MagmaLower) info = -1;
    else if (n < 0) info = -2;
    else if (nrhs < 0) info = -3;
    else if (ldda < std::max(1, n)) info = -5;
    else if (lddb < std::max(1, n)) info = -7;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return info;
    }
    hipStream_t st = stream_of(queue);
    if (n == 0 || nrhs == 0 || batchCount <= 0) {
        zero_info(dinfo_array, batchCount, st);
        return finish(MAGMA_SUCCESS, queue);
    }
    MatSet s;
    s.m = n; s.n = n; s.lda = ldda;
    s.A = dA_array; s.info = dinfo_array;
    s.B = dB_array; s.nrhs = nrhs; s.ldb = lddb;
    s.count = batchCount;
    if (uplo == MagmaUpper) {
        // U from the swapped lower factorization, then dpotrs('U'): U' solve,
        // U solve; a matrix with info > 0 keeps its right-hand sides
        magma_int_t r = launch_swap_triangles(s, st);
        if (r == MAGMA_SUCCESS) r = launch_potrf(s, n, false, st);
        if (r == MAGMA_SUCCESS) r = launch_swap_triangles(s, st);
        if (r == MAGMA_SUCCESS) r = launch_potrs(s, 'U', n, nrhs, st);
        return finish(r, queue);
    }
    return finish(launch_potrf(s, n, true, st), queue);
}

}  // extern "C"
