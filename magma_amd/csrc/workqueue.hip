// workqueue.hip -- device-side work lists for variable-size batches.
//
// The reference sizes vbatched grids by the maximum dimensions, which it reads
// back to the host (magmablas/vbatched_aux.cu:504-636, src/zpotrf_vbatched.cpp:350-351,
// magmablas/zgemm_vbatched.cpp:255-260), and lets blocks outside a matrix exit
// (magmablas/gemm_template_kernel_vbatched.cuh:53-54).  Here the lists are built on
// the device with no host round trip, so a vbatched call can be captured in a
// HIP graph:
//   * bucket_matrices: counting sort of matrix ids into size classes
//     (bucket b holds matrices whose size class is b), plus per-bucket counts;
//   * tile_prefix: exclusive prefix sum of per-matrix tile counts, so a
//     persistent kernel can map a flat tile index to (matrix, tile).
#include "workqueue.h"

namespace magma_amd {
namespace {

constexpr int kScanThreads = 1024;

// size class of an order-n matrix (see workqueue.h)
__device__ __forceinline__ int size_class(int m, int n)
{
    const int d = max(m, n);
    if (d <= 8) return 0;
    if (d <= 16) return 1;
    if (d <= 32) return 2;
    if (d <= 64) return 3;
    if (d <= 128) return 4;
    if (d <= 256) return 5;
    if (d <= 512) return 6;
    if (d <= 1024) return 7;
    if (d <= 2048) return 8;
    return 9;
}

__global__ void bucket_count_kernel(const int *m, const int *n, int count, int *counts)
{
    __shared__ int local[kNumBuckets];
    if (threadIdx.x < kNumBuckets) local[threadIdx.x] = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        const int mi = m ? m[i] : n[i];
        atomicAdd(&local[size_class(mi, n[i])], 1);
    }
    __syncthreads();
    if (threadIdx.x < kNumBuckets && local[threadIdx.x]) atomicAdd(&counts[threadIdx.x], local[threadIdx.x]);
}

// counts[b] -> offsets[b] (exclusive), cursor[b] = offsets[b]
__global__ void bucket_offsets_kernel(WorkQueue q)
{
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int b = 0; b < kNumBuckets; ++b) {
            q.offsets[b] = acc;
            q.cursor[b] = acc;
            acc += q.counts[b];
        }
        q.offsets[kNumBuckets] = acc;
    }
}

__global__ void bucket_scatter_kernel(const int *m, const int *n, int count, WorkQueue q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        const int mi = m ? m[i] : n[i];
        const int slot = atomicAdd(&q.cursor[size_class(mi, n[i])], 1);
        q.ids[slot] = i;
    }
}

__global__ void zero_kernel(int *p, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

// One workgroup: prefix[i] = sum_{j<i} tiles(j), prefix[count] = total.
__global__ void __launch_bounds__(kScanThreads) tile_prefix_kernel(const int *m, const int *n, int count,
                                                                  int tile_m, int tile_n, int *prefix)
{
    __shared__ int part[kScanThreads];
    const int tid = threadIdx.x;
    const int per = (count + kScanThreads - 1) / kScanThreads;
    const int lo = min(count, tid * per), hi = min(count, lo + per);
    int s = 0;
    for (int i = lo; i < hi; ++i) {
        const int mi = max(m[i], 0), ni = max(n[i], 0);
        s += ((mi + tile_m - 1) / tile_m) * ((ni + tile_n - 1) / tile_n);
    }
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int v = (tid >= off) ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int acc = part[tid] - s;                 // exclusive prefix of this thread's chunk
    for (int i = lo; i < hi; ++i) {
        prefix[i] = acc;
        const int mi = max(m[i], 0), ni = max(n[i], 0);
        acc += ((mi + tile_m - 1) / tile_m) * ((ni + tile_n - 1) / tile_n);
    }
    if (tid == kScanThreads - 1) prefix[count] = part[tid];
}

// One workgroup: the ids list[0 .. *count) sorted into size classes by n[id].
// Used for short lists built on the device (the FP64 fallbacks of the mixed
// solvers), so one workgroup is enough.
__global__ void __launch_bounds__(kScanThreads) bucket_list_kernel(const int *n, const int *list,
                                                                  const int *count, WorkQueue q)
{
    __shared__ int cnt[kNumBuckets];
    __shared__ int cur[kNumBuckets];
    const int tid = threadIdx.x;
    if (tid < kNumBuckets) cnt[tid] = 0;
    __syncthreads();
    const int c = *count;
    for (int i = tid; i < c; i += kScanThreads) atomicAdd(&cnt[size_class(n[list[i]], n[list[i]])], 1);
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int b = 0; b < kNumBuckets; ++b) {
            q.counts[b] = cnt[b];
            q.offsets[b] = acc;
            cur[b] = acc;
            acc += cnt[b];
        }
        q.offsets[kNumBuckets] = acc;
    }
    __syncthreads();
    for (int i = tid; i < c; i += kScanThreads) {
        const int id = list[i];
        q.ids[atomicAdd(&cur[size_class(n[id], n[id])], 1)] = id;
    }
}

}  // namespace

magma_int_t build_workqueue_list(const int *n, const int *list, const int *count, WorkQueue q, hipStream_t st)
{
    bucket_list_kernel<<<1, kScanThreads, 0, st>>>(n, list, count, q);
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

size_t workqueue_bytes(int count) { return sizeof(int) * (size_t)(count + 3 * kNumBuckets + 1); }

WorkQueue workqueue_view(void *mem, int count)
{
    WorkQueue q;
    int *p = static_cast<int *>(mem);
    q.ids = p;
    q.counts = p + count;
    q.offsets = q.counts + kNumBuckets;
    q.cursor = q.offsets + kNumBuckets + 1;
    return q;
}

magma_int_t build_workqueue(const int *m, const int *n, int count, WorkQueue q, hipStream_t st)
{
    zero_kernel<<<1, 64, 0, st>>>(q.counts, kNumBuckets);
    bucket_count_kernel<<<ceil_div(count, 256), 256, 0, st>>>(m, n, count, q.counts);
    bucket_offsets_kernel<<<1, 64, 0, st>>>(q);
    bucket_scatter_kernel<<<ceil_div(count, 256), 256, 0, st>>>(m, n, count, q);
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

magma_int_t build_tile_prefix(const int *m, const int *n, int count, int tile_m, int tile_n, int *prefix,
                              hipStream_t st)
{
    tile_prefix_kernel<<<1, kScanThreads, 0, st>>>(m, n, count, tile_m, tile_n, prefix);
    MAGMA_LAUNCH_CHECK();
    return MAGMA_SUCCESS;
}

}  // namespace magma_amd
