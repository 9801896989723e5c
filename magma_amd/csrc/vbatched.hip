// vbatched.hip -- variable-size entry points and the batched GEMM.
//
// A vbatched call bins the matrices by size class on the device
// (workqueue.hip) and runs each class on the kernel family built for it:
// register-resident kernels up to 64, blocked workgroup-per-matrix kernels
// above.  Two modes, as in the reference's API:
//   * checked entry points (magma_dgetrf_vbatched, ...) validate the size
//     arrays with a device kernel and read the flags and per-class counts back
//     in one synchronizing copy (the reference reads back its checker flags
//     and max sizes the same way, magmablas/vbatched_check.cu:48-105,
//     src/zpotrf_vbatched.cpp:343-351); each non-empty class is then launched
//     with its exact count;
//   * *_max_nocheck entry points skip validation and never synchronize: the
//     class kernels are persistent and read their counts on the device, so the
//     whole call can be captured in a HIP graph.
#include <algorithm>
#include <cstdio>
#include <mutex>

#include "kernels.h"
#include "mixed.h"
#include "workqueue.h"
#include "tune.h"

// Work list of magma_dgetrf_vbatched_max_nocheck_noalloc: the size-class sort
// of up to kNoAllocMaxBatch matrix ids (workqueue_view layout), in a
// file-scope device array so that the call allocates nothing.  Each device
// holds its own copy (4 MiB) until the process ends.  In host code `scratch`
// is only a symbol: its device address comes from hipGetSymbolAddress.
namespace {
constexpr int kNoAllocMaxBatch = 1 << 20;
constexpr int kNoAllocMaxDevices = 64;
}  // namespace
static __device__ __align__(256) int scratch[kNoAllocMaxBatch + 3 * magma_amd::kNumBuckets + 1];

namespace magma_amd {

magma_int_t launch_getrf_small_v(const MatSet &s, int max_mn_dim, bool solve, bool var, hipStream_t st);
magma_int_t launch_potrf_small_v(const MatSet &s, int max_n, bool solve, bool var, hipStream_t st);
magma_int_t launch_getrf_blocked(const MatSet &s, int max_m, hipStream_t st);
magma_int_t launch_getrf_mfma(const MatSet &s, int max_dim, hipStream_t st);
magma_int_t launch_potrf_blocked(const MatSet &s, int max_n, hipStream_t st);
magma_int_t launch_gemm_tiles(const GemmSet &g, const int *prefix, long max_tiles, int tile, hipStream_t st);
magma_int_t zero_info(int *info, int count, hipStream_t st);

namespace {

// flag slots written by the checkers
enum { kFlagM = 0, kFlagN, kFlagK, kFlagLda, kFlagLdb, kFlagLdc, kFlagNrhs, kFlagTooBig, kNumFlags };

__global__ void clear_kernel(int *p, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

// Validates per-matrix sizes.  mode: 'F' factorization (m, n, lda >= max(1,m)),
// 'C' Cholesky (n, lda >= n, reference rule), 'S' solve (n, nrhs, lda, ldb),
// 'M' mixed solve (n, nrhs, lda, ldb, ldx = ldc), 'G' gemm (m, n, k, lda,
// ldb, ldc for the given transposes).
__global__ void check_kernel(char mode, char ta, char tb, const int *m, const int *n, const int *k,
                             const int *lda, const int *ldb, const int *ldc, int count, int limit,
                             int *flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int ni = n[i];
    const int mi = m ? m[i] : ni;
    if (mode == 'F') {
        if (mi < 0) flags[kFlagM] = 1;
        if (ni < 0) flags[kFlagN] = 1;
        if (lda[i] < max(1, mi)) flags[kFlagLda] = 1;
        if (max(mi, ni) > limit) flags[kFlagTooBig] = 1;
    } else if (mode == 'C') {
        if (ni < 0) flags[kFlagN] = 1;
        if (lda[i] < ni) flags[kFlagLda] = 1;
        if (ni > limit) flags[kFlagTooBig] = 1;
    } else if (mode == 'S' || mode == 'M') {
        if (ni < 0) flags[kFlagN] = 1;
        if (k[i] < 0) flags[kFlagNrhs] = 1;
        if (lda[i] < max(1, ni)) flags[kFlagLda] = 1;
        if (ldb[i] < max(1, ni)) flags[kFlagLdb] = 1;
        if (mode == 'M' && ldc[i] < max(1, ni)) flags[kFlagLdc] = 1;
        if (ni > limit) flags[kFlagTooBig] = 1;
    } else {
        const int ki = k[i];
        if (mi < 0) flags[kFlagM] = 1;
        if (ni < 0) flags[kFlagN] = 1;
        if (ki < 0) flags[kFlagK] = 1;
        if (ta == 'N' ? lda[i] < mi : lda[i] < ki) flags[kFlagLda] = 1;
        if (tb == 'N' ? ldb[i] < ki : ldb[i] < ni) flags[kFlagLdb] = 1;
        if (ldc[i] < mi) flags[kFlagLdc] = 1;
    }
}

struct VScratch {
    int *flags = nullptr;
    WorkQueue q;
    int *prefix = nullptr;      // gemm tile prefix [count + 1]
    int *host = nullptr;        // pinned: flags + counts
};

// Carves the queue workspace into checker flags, work queue and tile prefix.
bool vscratch(magma_queue_t queue, int count, VScratch &v)
{
    const size_t fl = round_up(sizeof(int) * kNumFlags, 256);
    const size_t wq = round_up(workqueue_bytes(count), 256);
    const size_t pf = round_up(sizeof(int) * (size_t)(count + 1), 256);
    char *base = static_cast<char *>(queue_workspace(queue, fl + wq + pf));
    if (!base) return false;
    v.flags = reinterpret_cast<int *>(base);
    v.q = workqueue_view(base + fl, count);
    v.prefix = reinterpret_cast<int *>(base + fl + wq);
    v.host = static_cast<int *>(queue_host(queue, sizeof(int) * (kNumFlags + 2 * kNumBuckets)));
    return v.host != nullptr;
}

// Runs the checker, builds the work queue and reads flags + class counts back.
magma_int_t check_and_bucket(magma_queue_t queue, VScratch &v, char mode, char ta, char tb,
                             const int *m, const int *n, const int *k, const int *lda, const int *ldb,
                             const int *ldc, int count, int limit, bool bucket)
{
    hipStream_t st = stream_of(queue);
    clear_kernel<<<1, 64, 0, st>>>(v.flags, kNumFlags);
    check_kernel<<<ceil_div(count, 256), 256, 0, st>>>(mode, ta, tb, m, n, k, lda, ldb, ldc, count, limit,
                                                       v.flags);
    MAGMA_LAUNCH_CHECK();
    if (bucket) {
        magma_int_t r = build_workqueue(m, n, count, v.q, st);
        if (r != MAGMA_SUCCESS) return r;
    }
    MAGMA_HIP_CHECK(hipMemcpyAsync(v.host, v.flags, sizeof(int) * kNumFlags, hipMemcpyDeviceToHost, st));
    if (bucket) {
        MAGMA_HIP_CHECK(hipMemcpyAsync(v.host + kNumFlags, v.q.counts, sizeof(int) * kNumBuckets,
                                       hipMemcpyDeviceToHost, st));
        MAGMA_HIP_CHECK(hipMemcpyAsync(v.host + kNumFlags + kNumBuckets, v.q.offsets, sizeof(int) * kNumBuckets,
                                       hipMemcpyDeviceToHost, st));
    }
    MAGMA_HIP_CHECK(hipStreamSynchronize(st));
    return MAGMA_SUCCESS;
}

magma_int_t finish(magma_int_t r, magma_queue_t queue)
{
    if (r == MAGMA_SUCCESS && !(queue && queue->async)) {
        hipError_t e = hipStreamSynchronize(stream_of(queue));
        if (e != hipSuccess) return report_hip_error(e, "hipStreamSynchronize", __FILE__, __LINE__);
    }
    return r;
}

// Launches one size class.  kind: 'L' getrf, 'C' potrf, 'N'/'T' getrs, 'l'/'u' potrs.
magma_int_t launch_class(char kind, int b, MatSet s, hipStream_t st)
{
    const int lim = kBucketLimit[b];
    switch (kind) {
    case 'L':
        if (lim <= 64) return launch_getrf_small_v(s, lim, false, true, st);
        return lim <= 512 ? launch_getrf_mfma(s, lim, st) : launch_getrf_blocked(s, lim, st);
    case 'C': return lim <= 64 ? launch_potrf_small_v(s, lim, false, true, st) : launch_potrf_blocked(s, lim, st);
    case 'N':
    case 'T': return launch_getrs(s, kind, lim, 0, st);
    case 'l': return launch_potrs(s, 'L', lim, 0, st);
    default:  return launch_potrs(s, 'U', lim, 0, st);
    }
}

// Checked path: exact per-class counts from the host copy.
magma_int_t run_classes(char kind, MatSet s, const VScratch &v, hipStream_t st)
{
    const int *counts = v.host + kNumFlags;
    const int *offsets = v.host + kNumFlags + kNumBuckets;
    for (int b = 0; b < kNumBuckets; ++b) {
        if (counts[b] == 0) continue;
        MatSet c = s;
        c.index = v.q.ids + offsets[b];
        c.count = counts[b];
        magma_int_t r = launch_class(kind, b, c, st);
        if (r != MAGMA_SUCCESS) return r;
    }
    return MAGMA_SUCCESS;
}

// Unchecked path: persistent class kernels reading their counts on the device,
// for every class up to the one holding max_dim.
magma_int_t run_classes_device(char kind, MatSet s, const VScratch &v, int max_dim, hipStream_t st)
{
    for (int b = 0; b < kNumBuckets; ++b) {
        if (b > 0 && kBucketLimit[b - 1] >= max_dim) break;
        MatSet c = s;
        c.index = v.q.ids;
        c.dev_offset = v.q.offsets + b;
        c.dev_count = v.q.counts + b;
        magma_int_t r = launch_class(kind, b, c, st);
        if (r != MAGMA_SUCCESS) return r;
    }
    return MAGMA_SUCCESS;
}

// Largest order of the factor kernels (above 2048 the blocked kernels narrow
// their panels to keep the panel column in registers).
constexpr int kMaxFactorDim = 8192;

// Appends ids[0 .. cnt) to a fallback list with dsgesv's iter code -1 (the
// order is beyond the mixed kernels).
__global__ void append_fallback_kernel(const int *ids, int cnt, int *list, int *count, int *iter)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) {
        const int id = ids[i];
        list[atomicAdd(count, 1)] = id;
        if (iter) iter[id] = -1;
    }
}

// Mixed-precision solves of matrices of different orders (kind 'L' LU, 'C'
// Cholesky).  After the checker's read-back the host knows how many matrices
// each size class holds:
//   * classes up to 256 form one contiguous range of the class-sorted ids and
//     take one persistent launch of the n <= 256 mixed kernel, the 512 class
//     one launch of the n <= 512 kernel (larger systems first);
//   * orders above 512 go straight to the FP64 list, with iter = -1;
//   * the FP64 list (those plus the matrices whose refinement gave up) is
//     binned by class on the device and every class present in the input gets
//     its FP64 factor and solve kernels, which read their counts on the device.
magma_int_t mixed_vbatched(char kind, bool upper, int *n, int *nrhs, double **dA, int *ldda, int **dipiv,
                           double **dB, int *lddb, double **dX, int *lddx, int *info, int *iter, int count,
                           magma_queue_t queue, const char *name)
{
    hipStream_t st = stream_of(queue);
    const size_t fl = round_up(sizeof(int) * kNumFlags, 256);
    const size_t wq = round_up(workqueue_bytes(count), 256);
    const size_t lb = round_up(sizeof(int) * (size_t)(count + 1), 256);
    const size_t head = fl + wq + lb;
    char *base = static_cast<char *>(queue_workspace(queue, head));
    VScratch v;
    v.host = static_cast<int *>(queue_host(queue, sizeof(int) * (kNumFlags + 2 * kNumBuckets)));
    if (!base || !v.host) return MAGMA_ERR_DEVICE_ALLOC;
    v.flags = reinterpret_cast<int *>(base);
    v.q = workqueue_view(base + fl, count);
    magma_int_t r = check_and_bucket(queue, v, 'M', 'N', 'N', nullptr, n, nrhs, ldda, lddb, lddx, count,
                                     kMaxFactorDim, true);
    if (r != MAGMA_SUCCESS) return r;
    const int *f = v.host;
    magma_int_t arg = 0;
    if (f[kFlagN]) arg = -1;
    else if (f[kFlagNrhs]) arg = -2;
    else if (f[kFlagLda]) arg = -4;
    else if (f[kFlagLdb]) arg = -7;
    else if (f[kFlagLdc]) arg = -9;
    if (arg != 0) {
        magma_xerbla(name, -arg);
        return arg;
    }
    if (f[kFlagTooBig]) return MAGMA_ERR_NOT_SUPPORTED;
    // Upper: the lower kernels on the swapped triangles, swapped back at the end
    MatSet tri;
    tri.m_arr = n; tri.n_arr = n; tri.lda_arr = ldda;
    tri.A = dA;
    tri.count = count;
    if (upper) {
        r = launch_swap_triangles(tri, st);
        if (r != MAGMA_SUCCESS) return r;
    }
    int counts[kNumBuckets], offsets[kNumBuckets];
    std::copy(v.host + kNumFlags, v.host + kNumFlags + kNumBuckets, counts);
    std::copy(v.host + kNumFlags + kNumBuckets, v.host + kNumFlags + 2 * kNumBuckets, offsets);
    int n_small = 0, max_small = 0, max_all = 0;
    for (int b = 0; b < kNumBuckets; ++b) {
        if (counts[b] == 0) continue;
        max_all = kBucketLimit[b];
        if (kBucketLimit[b] <= 256) {
            n_small += counts[b];
            max_small = kBucketLimit[b];
        }
    }
    const int n_mid = counts[6];
    const int n_large = counts[7] + counts[8] + counts[9];
    // factor slots and status entries: each launch has its own plan; the
    // launches are ordered on the stream, so they share one region
    size_t slot_bytes = 0;
    const MixedPlan plan_small = n_small ? mixed_plan(n_small, max_small, 0) : MixedPlan();
    const MixedPlan plan_mid = n_mid ? mixed_plan(n_mid, 512, 0) : MixedPlan();
    if (n_small) slot_bytes = mixed_workspace_bytes(plan_small, max_small, true);
    if (n_mid) slot_bytes = std::max(slot_bytes, mixed_workspace_bytes(plan_mid, 512, true));
    char *ws = static_cast<char *>(queue_workspace(queue, head + round_up(slot_bytes, 256)));
    if (!ws) return MAGMA_ERR_DEVICE_ALLOC;
    v.q = workqueue_view(ws + fl, count);
    if (ws != base) {
        // the workspace moved (and lost the lists): bin again; the stream order
        // makes the new lists those the launches below read
        r = build_workqueue(nullptr, n, count, v.q, st);
        if (r != MAGMA_SUCCESS) return r;
    }
    int *fb_count = reinterpret_cast<int *>(ws + fl + wq);
    int *fb_list = fb_count + 1;
    r = launch_zero_counts(fb_count, nullptr, st);
    if (r != MAGMA_SUCCESS) return r;
    MixedSet g;
    g.n_arr = n; g.nrhs_arr = nrhs; g.lda_arr = ldda; g.ldb_arr = lddb; g.ldx_arr = lddx;
    g.kind = kind;
    g.A = dA; g.B = dB; g.X = dX;
    g.ipiv = dipiv; g.info = info; g.iter = iter;
    if (kind == 'C') {
        g.fbc_list = fb_list; g.fbc_count = fb_count;
    } else {
        g.fb_list = fb_list; g.fb_count = fb_count;
    }
    if (n_mid) {
        MixedSet gm = g;
        gm.index = v.q.ids + offsets[6];
        gm.count = n_mid;
        mixed_workspace_view(gm, plan_mid, 512, true, ws + head);
        r = launch_mixed(gm, 512, plan_mid, st);
        if (r != MAGMA_SUCCESS) return r;
    }
    if (n_small) {
        MixedSet gs = g;
        gs.index = v.q.ids + offsets[0];
        gs.count = n_small;
        mixed_workspace_view(gs, plan_small, max_small, true, ws + head);
        r = launch_mixed(gs, max_small, plan_small, st);
        if (r != MAGMA_SUCCESS) return r;
    }
    if (n_large) {
        append_fallback_kernel<<<ceil_div(n_large, 256), 256, 0, st>>>(v.q.ids + offsets[7], n_large, fb_list,
                                                                       fb_count, iter);
        MAGMA_LAUNCH_CHECK();
    }
    // FP64 path over the list: X <- B, then factor and solve in place, class
    // by class (the class lists overwrite the input's, which the launches
    // above have finished with by stream order)
    g.count = count;
    r = launch_copy_rhs(g, fb_list, fb_count, st);
    if (r != MAGMA_SUCCESS) return r;
    r = build_workqueue_list(n, fb_list, fb_count, v.q, st);
    if (r != MAGMA_SUCCESS) return r;
    MatSet s;
    s.m_arr = n; s.n_arr = n; s.lda_arr = ldda;
    s.A = dA; s.ipiv = dipiv; s.info = info;
    s.B = dX; s.nrhs_arr = nrhs; s.ldb_arr = lddx;
    s.count = count;
    for (int pass = 0; pass < 2; ++pass) {
        const char k = kind == 'C' ? (pass == 0 ? 'C' : 'l') : (pass == 0 ? 'L' : 'N');
        for (int b = 0; b < kNumBuckets && kBucketLimit[b] <= max_all; ++b) {
            if (counts[b] == 0) continue;
            MatSet c = s;
            c.index = v.q.ids;
            c.dev_offset = v.q.offsets + b;
            c.dev_count = v.q.counts + b;
            c.count = counts[b];
            r = launch_class(k, b, c, st);
            if (r != MAGMA_SUCCESS) return r;
        }
    }
    if (upper) {
        r = launch_swap_triangles(tri, st);
        if (r != MAGMA_SUCCESS) return r;
    }
    return finish(MAGMA_SUCCESS, queue);
}

}  // namespace
}  // namespace magma_amd

using namespace magma_amd;

extern "C" {

// ---------------------------------------------------------------- LU -------
magma_int_t magma_dgetrf_vbatched(magma_int_t *m, magma_int_t *n, double **dA_array, magma_int_t *ldda,
                                  magma_int_t **dipiv_array, magma_int_t *info_array,
                                  magma_int_t batchCount, magma_queue_t queue)
{
    if (batchCount < 0) {
        magma_xerbla(__func__, 8);
        return -8;
    }
    if (batchCount == 0) return MAGMA_SUCCESS;
    VScratch v;
    if (!vscratch(queue, batchCount, v)) return MAGMA_ERR_DEVICE_ALLOC;
    magma_int_t r = check_and_bucket(queue, v, 'F', 'N', 'N', m, n, nullptr, ldda, nullptr, nullptr,
                                     batchCount, kMaxFactorDim, true);
    if (r != MAGMA_SUCCESS) return r;
    const int *f = v.host;
    magma_int_t arg = 0;
    if (f[kFlagM]) arg = -1;
    else if (f[kFlagN]) arg = -2;
    else if (f[kFlagLda]) arg = -4;
    if (arg != 0) {
        zero_info(info_array, batchCount, stream_of(queue));
        magma_xerbla(__func__, -arg);
        return arg;
    }
    if (f[kFlagTooBig]) return MAGMA_ERR_NOT_SUPPORTED;
    MatSet s;
    s.m_arr = m; s.n_arr = n; s.lda_arr = ldda;
    s.A = dA_array; s.ipiv = dipiv_array; s.info = info_array;
    return finish(run_classes('L', s, v, stream_of(queue)), queue);
}

magma_int_t magma_dgetrf_vbatched_max_nocheck(magma_int_t *m, magma_int_t *n, magma_int_t max_mn,
                                              double **dA_array, magma_int_t *ldda,
                                              magma_int_t **dipiv_array, magma_int_t *info_array,
                                              magma_int_t batchCount, magma_queue_t queue)
{
    if (batchCount <= 0) return MAGMA_SUCCESS;
    if (max_mn > kMaxFactorDim) return MAGMA_ERR_NOT_SUPPORTED;
    VScratch v;
    if (!vscratch(queue, batchCount, v)) return MAGMA_ERR_DEVICE_ALLOC;
    hipStream_t st = stream_of(queue);
    magma_int_t r = build_workqueue(m, n, batchCount, v.q, st);
    if (r != MAGMA_SUCCESS) return r;
    MatSet s;
    s.m_arr = m; s.n_arr = n; s.lda_arr = ldda;
    s.A = dA_array; s.ipiv = dipiv_array; s.info = info_array;
    s.count = batchCount;
    return finish(run_classes_device('L', s, v, max_mn, st), queue);
}

}  // extern "C"

namespace {

// Device address of `scratch` on each device, resolved once per device with
// that device current (every device has its own copy at its own address).
std::mutex g_scratch_mu;
void *g_scratch_ptr[kNoAllocMaxDevices];
hipError_t g_scratch_err[kNoAllocMaxDevices];
bool g_scratch_done[kNoAllocMaxDevices];

magma_int_t scratch_address(int device, void **out)
{
    *out = nullptr;
    if (device < 0 || device >= kNoAllocMaxDevices) return MAGMA_ERR_ILLEGAL_VALUE;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_done[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != device) e = hipSetDevice(device);
        void *p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
        g_scratch_ptr[device] = (e == hipSuccess) ? p : nullptr;
        g_scratch_err[device] = e;
        g_scratch_done[device] = true;
    }
    if (g_scratch_err[device] != hipSuccess)
        return report_hip_error(g_scratch_err[device], "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    *out = g_scratch_ptr[device];
    return MAGMA_SUCCESS;
}

}  // namespace

extern "C" {

// magma_dgetrf_vbatched_max_nocheck with its work list in the file-scope
// `scratch` array instead of the queue's workspace: no allocation of any kind,
// for callers that cannot allocate.  Calls that overlap on one device share
// the array and race; the kernels rebuild the whole list on every call
// (counts cleared by a kernel), so nothing carries over between calls.
magma_int_t magma_dgetrf_vbatched_max_nocheck_noalloc(magma_int_t *m, magma_int_t *n, magma_int_t max_mn,
                                                      double **dA_array, magma_int_t *ldda,
                                                      magma_int_t **dipiv_array, magma_int_t *info_array,
                                                      magma_int_t batchCount, magma_queue_t queue)
{
    if (batchCount <= 0) return MAGMA_SUCCESS;
    if (batchCount > kNoAllocMaxBatch) {
        std::fprintf(stderr, "%s: batchCount %d exceeds the %d matrices of its static work list\n", __func__,
                     (int)batchCount, kNoAllocMaxBatch);
        magma_xerbla(__func__, 8);
        return -8;
    }
    if (max_mn > kMaxFactorDim) return MAGMA_ERR_NOT_SUPPORTED;
    void *base = nullptr;
    const magma_int_t e = scratch_address(queue ? queue->device : 0, &base);
    if (e != MAGMA_SUCCESS) return e;
    VScratch v;
    v.q = workqueue_view(base, batchCount);
    hipStream_t st = stream_of(queue);
    magma_int_t r = build_workqueue(m, n, batchCount, v.q, st);
    if (r != MAGMA_SUCCESS) return r;
    MatSet s;
    s.m_arr = m; s.n_arr = n; s.lda_arr = ldda;
    s.A = dA_array; s.ipiv = dipiv_array; s.info = info_array;
    s.count = batchCount;
    return finish(run_classes_device('L', s, v, max_mn, st), queue);
}

magma_int_t magma_dgetrs_vbatched(magma_trans_t trans, magma_int_t *n, magma_int_t *nrhs,
                                  double **dA_array, magma_int_t *ldda, magma_int_t **dipiv_array,
                                  double **dB_array, magma_int_t *lddb,
                                  magma_int_t batchCount, magma_queue_t queue)
{
    if (trans != MagmaNoTrans && trans != MagmaTrans && trans != MagmaConjTrans) {
        magma_xerbla(__func__, 1);
        return -1;
    }
    if (batchCount < 0) {
        magma_xerbla(__func__, 10);
        return -10;
    }
    if (batchCount == 0) return MAGMA_SUCCESS;
    VScratch v;
    if (!vscratch(queue, batchCount, v)) return MAGMA_ERR_DEVICE_ALLOC;
    magma_int_t r = check_and_bucket(queue, v, 'S', 'N', 'N', nullptr, n, nrhs, ldda, lddb, nullptr,
                                     batchCount, kMaxFactorDim, true);
    if (r != MAGMA_SUCCESS) return r;
    const int *f = v.host;
    magma_int_t arg = 0;
    if (f[kFlagN]) arg = -2;
    else if (f[kFlagNrhs]) arg = -3;
    else if (f[kFlagLda]) arg = -5;
    else if (f[kFlagLdb]) arg = -8;
    if (arg != 0) {
        magma_xerbla(__func__, -arg);
        return arg;
    }
    if (f[kFlagTooBig]) return MAGMA_ERR_NOT_SUPPORTED;
    MatSet s;
    s.n_arr = n; s.m_arr = n; s.lda_arr = ldda;
    s.nrhs_arr = nrhs; s.ldb_arr = lddb;
    s.A = dA_array; s.ipiv = dipiv_array; s.B = dB_array;
    return finish(run_classes(trans == MagmaNoTrans ? 'N' : 'T', s, v, stream_of(queue)), queue);
}

// Unchecked, never synchronizing (capturable in a HIP graph): max_n bounds
// every n[i]; nrhs[i] and lddb[i] are not validated.
magma_int_t magma_dgetrs_vbatched_max_nocheck(magma_trans_t trans, magma_int_t *n, magma_int_t *nrhs,
                                              double **dA_array, magma_int_t *ldda, magma_int_t **dipiv_array,
                                              double **dB_array, magma_int_t *lddb,
                                              magma_int_t batchCount, magma_int_t max_n, magma_queue_t queue)
{
    if (trans != MagmaNoTrans && trans != MagmaTrans && trans != MagmaConjTrans) {
        magma_xerbla(__func__, 1);
        return -1;
    }
    if (batchCount <= 0) return MAGMA_SUCCESS;
    if (max_n > kMaxFactorDim) return MAGMA_ERR_NOT_SUPPORTED;
    VScratch v;
    if (!vscratch(queue, batchCount, v)) return MAGMA_ERR_DEVICE_ALLOC;
    hipStream_t st = stream_of(queue);
    magma_int_t r = build_workqueue(nullptr, n, batchCount, v.q, st);
    if (r != MAGMA_SUCCESS) return r;
    MatSet s;
    s.n_arr = n; s.m_arr = n; s.lda_arr = ldda;
    s.nrhs_arr = nrhs; s.ldb_arr = lddb;
    s.A = dA_array; s.ipiv = dipiv_array; s.B = dB_array;
    s.count = batchCount;
    return finish(run_classes_device(trans == MagmaNoTrans ? 'N' : 'T', s, v, max_n, st), queue);
}

// ---------------------------------------------------------------- Cholesky ---
// src/zpotrf_vbatched.cpp:335-356; checks magmablas/vbatched_check.cu:840-892
magma_int_t magma_dpotrf_vbatched(magma_uplo_t uplo, magma_int_t *n, double **dA_array, magma_int_t *ldda,
                                  magma_int_t *info_array, magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t arg = 0;
    if (uplo != MagmaLower && uplo != MagmaUpper) arg = -1;
    else if (batchCount < 0) arg = -6;
    if (arg != 0) {
        magma_xerbla(__func__, -arg);
        return arg;
    }
    if (batchCount == 0) return MAGMA_SUCCESS;
    VScratch v;
    if (!vscratch(queue, batchCount, v)) return MAGMA_ERR_DEVICE_ALLOC;
    magma_int_t r = check_and_bucket(queue, v, 'C', 'N', 'N', nullptr, n, nullptr, ldda, nullptr, nullptr,
                                     batchCount, kMaxFactorDim, true);
    if (r != MAGMA_SUCCESS) return r;
    const int *f = v.host;
    if (f[kFlagN]) arg = -2;
    else if (f[kFlagLda]) arg = -4;
    if (arg != 0) {
        magma_xerbla(__func__, -arg);
        return arg;
    }
    if (f[kFlagTooBig]) return MAGMA_ERR_NOT_SUPPORTED;
    MatSet s;
    s.n_arr = n; s.m_arr = n; s.lda_arr = ldda;
    s.A = dA_array; s.info = info_array;
    if (uplo == MagmaUpper) {
        // the lower kernels between two in-place triangle swaps (uplo.hip)
        MatSet tri = s;
        tri.count = batchCount;
        magma_int_t e = launch_swap_triangles(tri, stream_of(queue));
        if (e == MAGMA_SUCCESS) e = run_classes('C', s, v, stream_of(queue));
        if (e == MAGMA_SUCCESS) e = launch_swap_triangles(tri, stream_of(queue));
        return finish(e, queue);
    }
    return finish(run_classes('C', s, v, stream_of(queue)), queue);
}

magma_int_t magma_dpotrf_vbatched_max_nocheck(magma_uplo_t uplo, magma_int_t *n, double **dA_array,
                                              magma_int_t *ldda, magma_int_t *info_array,
                                              magma_int_t batchCount, magma_int_t max_n, magma_queue_t queue)
{
    if (uplo != MagmaLower && uplo != MagmaUpper) {
        magma_xerbla(__func__, 1);
        return -1;
    }
    if (batchCount <= 0) return MAGMA_SUCCESS;
    if (max_n > kMaxFactorDim) return MAGMA_ERR_NOT_SUPPORTED;
    VScratch v;
    if (!vscratch(queue, batchCount, v)) return MAGMA_ERR_DEVICE_ALLOC;
    hipStream_t st = stream_of(queue);
    magma_int_t r = build_workqueue(nullptr, n, batchCount, v.q, st);
    if (r != MAGMA_SUCCESS) return r;
    MatSet s;
    s.n_arr = n; s.m_arr = n; s.lda_arr = ldda;
    s.A = dA_array; s.info = info_array;
    s.count = batchCount;
    if (uplo == MagmaUpper) {
        r = launch_swap_triangles(s, st);
        if (r == MAGMA_SUCCESS) r = run_classes_device('C', s, v, max_n, st);
        if (r == MAGMA_SUCCESS) r = launch_swap_triangles(s, st);
        return finish(r, queue);
    }
    return finish(run_classes_device('C', s, v, max_n, st), queue);
}

magma_int_t magma_dpotrs_vbatched(magma_uplo_t uplo, magma_int_t *n, magma_int_t *nrhs,
                                  double **dA_array, magma_int_t *ldda, double **dB_array, magma_int_t *lddb,
                                  magma_int_t batchCount, magma_queue_t queue)
{
    if (uplo != MagmaLower && uplo != MagmaUpper) {
        magma_xerbla(__func__, 1);
        return -1;
    }
    if (batchCount < 0) {
        magma_xerbla(__func__, 9);
        return -9;
    }
    if (batchCount == 0) return MAGMA_SUCCESS;
    VScratch v;
    if (!vscratch(queue, batchCount, v)) return MAGMA_ERR_DEVICE_ALLOC;
    magma_int_t r = check_and_bucket(queue, v, 'S', 'N', 'N', nullptr, n, nrhs, ldda, lddb, nullptr,
                                     batchCount, kMaxFactorDim, true);
    if (r != MAGMA_SUCCESS) return r;
    const int *f = v.host;
    magma_int_t arg = 0;
    if (f[kFlagN]) arg = -2;
    else if (f[kFlagNrhs]) arg = -3;
    else if (f[kFlagLda]) arg = -5;
    else if (f[kFlagLdb]) arg = -7;
    if (arg != 0) {
        magma_xerbla(__func__, -arg);
        return arg;
    }
    if (f[kFlagTooBig]) return MAGMA_ERR_NOT_SUPPORTED;
    MatSet s;
    s.n_arr = n; s.m_arr = n; s.lda_arr = ldda;
    s.nrhs_arr = nrhs; s.ldb_arr = lddb;
    s.A = dA_array; s.B = dB_array;
    return finish(run_classes(uplo == MagmaLower ? 'l' : 'u', s, v, stream_of(queue)), queue);
}

// Unchecked, never synchronizing (capturable in a HIP graph).
magma_int_t magma_dpotrs_vbatched_max_nocheck(magma_uplo_t uplo, magma_int_t *n, magma_int_t *nrhs,
                                              double **dA_array, magma_int_t *ldda, double **dB_array,
                                              magma_int_t *lddb, magma_int_t batchCount, magma_int_t max_n,
                                              magma_queue_t queue)
{
    if (uplo != MagmaLower && uplo != MagmaUpper) {
        magma_xerbla(__func__, 1);
        return -1;
    }
    if (batchCount <= 0) return MAGMA_SUCCESS;
    if (max_n > kMaxFactorDim) return MAGMA_ERR_NOT_SUPPORTED;
    VScratch v;
    if (!vscratch(queue, batchCount, v)) return MAGMA_ERR_DEVICE_ALLOC;
    hipStream_t st = stream_of(queue);
    magma_int_t r = build_workqueue(nullptr, n, batchCount, v.q, st);
    if (r != MAGMA_SUCCESS) return r;
    MatSet s;
    s.n_arr = n; s.m_arr = n; s.lda_arr = ldda;
    s.nrhs_arr = nrhs; s.ldb_arr = lddb;
    s.A = dA_array; s.B = dB_array;
    s.count = batchCount;
    return finish(run_classes_device(uplo == MagmaLower ? 'l' : 'u', s, v, max_n, st), queue);
}

// ---------------------------------------------------------------- GEMM -------
static char tr(magma_trans_t t) { return t == MagmaNoTrans ? 'N' : 'T'; }

// magmablas/dgemm_batched_core.cu:180-200
void magmablas_dgemm_batched(magma_trans_t transA, magma_trans_t transB,
                             magma_int_t m, magma_int_t n, magma_int_t k, double alpha,
                             double const *const *dA_array, magma_int_t ldda,
                             double const *const *dB_array, magma_int_t lddb, double beta,
                             double **dC_array, magma_int_t lddc, magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t info = 0;
    if (transA != MagmaNoTrans && transA != MagmaTrans && transA != MagmaConjTrans) info = -1;
    else if (transB != MagmaNoTrans && transB != MagmaTrans && transB != MagmaConjTrans) info = -2;
    else if (m < 0) info = -3;
    else if (n < 0) info = -4;
    else if (k < 0) info = -5;
    else if (transA == MagmaNoTrans ? ldda < m : ldda < k) info = -8;
    else if (transB == MagmaNoTrans ? lddb < k : lddb < n) info = -10;
    else if (lddc < m) info = -13;
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return;
    }
    if (m == 0 || n == 0 || batchCount <= 0) return;
    if ((k == 0 || alpha == 0.0) && beta == 1.0) return;
    GemmSet g;
    g.ta = tr(transA); g.tb = tr(transB);
    g.m = m; g.n = n; g.k = k; g.lda = ldda; g.ldb = lddb; g.ldc = lddc;
    g.alpha = alpha; g.beta = beta;
    g.A = dA_array; g.B = dB_array; g.C = dC_array;
    g.count = batchCount;
    const int tile = gemm_tile_of(tuned_variant(kTuneGemm, size_class_of(std::max(m, n))));
    const long tiles = (long)ceil_div(m, tile) * ceil_div(n, tile) * batchCount;
    launch_gemm_tiles(g, nullptr, tiles, tile, stream_of(queue));
}

// magmablas/zgemm_batched.cpp:166-195 (no vendor-BLAS branch: one implementation)
void magma_dgemm_batched(magma_trans_t transA, magma_trans_t transB,
                         magma_int_t m, magma_int_t n, magma_int_t k, double alpha,
                         double const *const *dA_array, magma_int_t ldda,
                         double const *const *dB_array, magma_int_t lddb, double beta,
                         double **dC_array, magma_int_t lddc, magma_int_t batchCount, magma_queue_t queue)
{
    magmablas_dgemm_batched(transA, transB, m, n, k, alpha, dA_array, ldda, dB_array, lddb, beta,
                            dC_array, lddc, batchCount, queue);
}

void magmablas_dgemm_vbatched_max_nocheck(magma_trans_t transA, magma_trans_t transB,
                                          magma_int_t *m, magma_int_t *n, magma_int_t *k, double alpha,
                                          double const *const *dA_array, magma_int_t *ldda,
                                          double const *const *dB_array, magma_int_t *lddb, double beta,
                                          double **dC_array, magma_int_t *lddc, magma_int_t batchCount,
                                          magma_int_t max_m, magma_int_t max_n, magma_int_t max_k,
                                          magma_queue_t queue)
{
    (void)max_k;
    if (batchCount <= 0 || max_m <= 0 || max_n <= 0) return;
    VScratch v;
    if (!vscratch(queue, batchCount, v)) {
        std::fprintf(stderr, "%s: workspace allocation failed\n", __func__);
        return;
    }
    hipStream_t st = stream_of(queue);
    // one tile size for the whole call: the class of the largest matrix, or the
    // mixed-size class when the caller's bounds are not tight
    const int cls = (max_m > 2048 || max_n > 2048) ? kTuneClasses - 1 : size_class_of(std::max(max_m, max_n));
    const int tile = gemm_tile_of(tuned_variant(kTuneGemm, cls));
    if (build_tile_prefix(m, n, batchCount, tile, tile, v.prefix, st) != MAGMA_SUCCESS) return;
    GemmSet g;
    g.ta = tr(transA); g.tb = tr(transB);
    g.m_arr = m; g.n_arr = n; g.k_arr = k; g.lda_arr = ldda; g.ldb_arr = lddb; g.ldc_arr = lddc;
    g.alpha = alpha; g.beta = beta;
    g.A = dA_array; g.B = dB_array; g.C = dC_array;
    g.count = batchCount;
    const long tiles = (long)ceil_div(std::min(max_m, 4096), tile) * ceil_div(std::min(max_n, 4096), tile) * batchCount;
    launch_gemm_tiles(g, v.prefix, tiles, tile, st);
}

// magmablas/zgemm_vbatched.cpp:235-270
void magmablas_dgemm_vbatched(magma_trans_t transA, magma_trans_t transB,
                              magma_int_t *m, magma_int_t *n, magma_int_t *k, double alpha,
                              double const *const *dA_array, magma_int_t *ldda,
                              double const *const *dB_array, magma_int_t *lddb, double beta,
                              double **dC_array, magma_int_t *lddc,
                              magma_int_t batchCount, magma_queue_t queue)
{
    magma_int_t info = 0;
    if (transA != MagmaNoTrans && transA != MagmaTrans && transA != MagmaConjTrans) info = -1;
    else if (transB != MagmaNoTrans && transB != MagmaTrans && transB != MagmaConjTrans) info = -2;
    else if (batchCount < 0) info = -14;
    if (info == 0 && batchCount > 0) {
        VScratch v;
        if (!vscratch(queue, batchCount, v)) {
            std::fprintf(stderr, "%s: workspace allocation failed\n", __func__);
            return;
        }
        if (check_and_bucket(queue, v, 'G', tr(transA), tr(transB), m, n, k, ldda, lddb, lddc, batchCount,
                             1 << 30, false) != MAGMA_SUCCESS)
            return;
        const int *f = v.host;
        if (f[kFlagM]) info = -3;
        else if (f[kFlagN]) info = -4;
        else if (f[kFlagK]) info = -5;
        else if (f[kFlagLda]) info = -8;
        else if (f[kFlagLdb]) info = -10;
        else if (f[kFlagLdc]) info = -13;
    }
    if (info != 0) {
        magma_xerbla(__func__, -info);
        return;
    }
    if (batchCount == 0) return;
    // the tile list is exact (device prefix), so the max sizes only bound the grid
    magmablas_dgemm_vbatched_max_nocheck(transA, transB, m, n, k, alpha, dA_array, ldda, dB_array, lddb,
                                         beta, dC_array, lddc, batchCount, 1 << 12, 1 << 12, 1 << 12, queue);
}

void magmablas_dgemm_vbatched_max(magma_trans_t transA, magma_trans_t transB,
                                  magma_int_t *m, magma_int_t *n, magma_int_t *k, double alpha,
                                  double const *const *dA_array, magma_int_t *ldda,
                                  double const *const *dB_array, magma_int_t *lddb, double beta,
                                  double **dC_array, magma_int_t *lddc, magma_int_t batchCount,
                                  magma_int_t max_m, magma_int_t max_n, magma_int_t max_k,
                                  magma_queue_t queue)
{
    (void)max_m; (void)max_n; (void)max_k;
    magmablas_dgemm_vbatched(transA, transB, m, n, k, alpha, dA_array, ldda, dB_array, lddb, beta,
                             dC_array, lddc, batchCount, queue);
}

void magmablas_dgemm_vbatched_nocheck(magma_trans_t transA, magma_trans_t transB,
                                      magma_int_t *m, magma_int_t *n, magma_int_t *k, double alpha,
                                      double const *const *dA_array, magma_int_t *ldda,
                                      double const *const *dB_array, magma_int_t *lddb, double beta,
                                      double **dC_array, magma_int_t *lddc,
                                      magma_int_t batchCount, magma_queue_t queue)
{
    magmablas_dgemm_vbatched_max_nocheck(transA, transB, m, n, k, alpha, dA_array, ldda, dB_array, lddb,
                                         beta, dC_array, lddc, batchCount, 1 << 12, 1 << 12, 1 << 12, queue);
}

// ------------------------------------------------- mixed precision ---------
// No counterpart in the reference (it has no batched or vbatched dsgesv); the
// argument order and codes are those of magma_dsgesv_batched /
// magma_dsposv_batched, with per-matrix n, nrhs and leading dimensions in
// device arrays of batchCount+1 entries.
magma_int_t magma_dsgesv_vbatched(magma_int_t *n, magma_int_t *nrhs, double **dA_array, magma_int_t *ldda,
                                  magma_int_t **dipiv_array, double **dB_array, magma_int_t *lddb,
                                  double **dX_array, magma_int_t *lddx, magma_int_t *dinfo_array,
                                  magma_int_t *diter_array, magma_int_t batchCount, magma_queue_t queue)
{
    if (batchCount < 0) {
        magma_xerbla(__func__, 12);
        return -12;
    }
    if (batchCount == 0) return MAGMA_SUCCESS;
    return mixed_vbatched('L', false, n, nrhs, dA_array, ldda, dipiv_array, dB_array, lddb, dX_array, lddx,
                          dinfo_array, diter_array, batchCount, queue, __func__);
}

magma_int_t magma_dsposv_vbatched(magma_uplo_t uplo, magma_int_t *n, magma_int_t *nrhs, double **dA_array,
                                  magma_int_t *ldda, double **dB_array, magma_int_t *lddb,
                                  double **dX_array, magma_int_t *lddx, magma_int_t *dinfo_array,
                                  magma_int_t *diter_array, magma_int_t batchCount, magma_queue_t queue)
{
    if (uplo != MagmaLower && uplo != MagmaUpper) {
        magma_xerbla(__func__, 1);
        return -1;
    }
    if (batchCount < 0) {
        magma_xerbla(__func__, 12);
        return -12;
    }
    if (batchCount == 0) return MAGMA_SUCCESS;
    return mixed_vbatched('C', uplo == MagmaUpper, n, nrhs, dA_array, ldda, nullptr, dB_array, lddb, dX_array, lddx,
                          dinfo_array, diter_array, batchCount, queue, __func__);
}

}  // extern "C"
