// runtime.hip -- MAGMA runtime subset on HIP: init, devices, queues, memory,
// copies, events, errors, pointer-array helpers.
//
// Reference behaviour restated (not ported): interface_cuda/interface.cpp
// (init :168, queue create :690-716, sync :871), interface_cuda/alloc.cpp
// (malloc :60, free :93), interface_cuda/copy_v2.cpp, control/xerbla.cpp:51-70,
// interface_cuda/error.cpp, magmablas/zset_pointer.cu, magmablas/set_pointer.cu.
#include <sys/time.h>

#include <cstdlib>
#include <cstring>
#include <algorithm>

#include "kernels.h"

namespace magma_amd {

magma_int_t report_hip_error(hipError_t e, const char *what, const char *file, int line)
{
    std::fprintf(stderr, "libmagma_amd: HIP error '%s' (%d) in %s at %s:%d\n",
                 hipGetErrorString(e), (int)e, what, file, line);
    return MAGMA_ERR_UNKNOWN;
}

void *queue_workspace(magma_queue_t queue, size_t bytes)
{
    if (bytes <= queue->work_bytes) return queue->work;
    size_t want = round_up(bytes + bytes / 4, 1 << 20);
    if (queue->work) {
        (void)hipStreamSynchronize(queue->stream);
        (void)hipFree(queue->work);
        queue->work = nullptr;
        queue->work_bytes = 0;
    }
    if (hipMalloc(&queue->work, want) != hipSuccess) {
        queue->work = nullptr;
        return nullptr;
    }
    queue->work_bytes = want;
    return queue->work;
}

void *queue_host(magma_queue_t queue, size_t bytes)
{
    if (bytes <= queue->host_bytes) return queue->host;
    size_t want = round_up(bytes + bytes / 4, 1 << 16);
    if (queue->host) {
        (void)hipStreamSynchronize(queue->stream);
        (void)hipHostFree(queue->host);
        queue->host = nullptr;
        queue->host_bytes = 0;
    }
    if (hipHostMalloc(&queue->host, want, hipHostMallocDefault) != hipSuccess) {
        queue->host = nullptr;
        return nullptr;
    }
    queue->host_bytes = want;
    return queue->host;
}

int persistent_grid(long items, int per_cu)
{
    static int cus = 0;
    if (cus == 0) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
            cus = n;
        else
            cus = 256;
    }
    const long cap = (long)cus * per_cu;
    return (int)std::max(1L, std::min(items, cap));
}

}  // namespace magma_amd

using namespace magma_amd;

extern "C" {

magma_int_t magma_init(void)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return MAGMA_ERR_UNKNOWN;
    return MAGMA_SUCCESS;
}

magma_int_t magma_finalize(void) { return MAGMA_SUCCESS; }

void magma_version(magma_int_t *major, magma_int_t *minor, magma_int_t *micro)
{
    // API level of the reference this library is a drop-in for (ReleaseNotes:24).
    if (major) *major = 2;
    if (minor) *minor = 2;
    if (micro) *micro = 0;
}

real_Double_t magma_wtime(void)
{
    struct timeval t;
    gettimeofday(&t, nullptr);
    return t.tv_sec + t.tv_usec * 1e-6;
}

real_Double_t magma_sync_wtime(magma_queue_t queue)
{
    if (queue) hipStreamSynchronize(queue->stream);
    else hipDeviceSynchronize();
    return magma_wtime();
}

// control/xerbla.cpp:51-70
void magma_xerbla(const char *srname, magma_int_t neg_info)
{
    if (neg_info < 0)
        std::fprintf(stderr, "Error in %s, function-specific error (info = %d)\n", srname, -neg_info);
    else if (neg_info == 0)
        std::fprintf(stderr, "No error, why is %s calling xerbla? (info = %d)\n", srname, -neg_info);
    else if (neg_info >= -MAGMA_ERR)
        std::fprintf(stderr, "Error in %s, %s (info = %d)\n", srname, magma_strerror(-neg_info), -neg_info);
    else
        std::fprintf(stderr, "On entry to %s, parameter %d had an illegal value (info = %d)\n",
                     srname, neg_info, -neg_info);
}

const char *magma_strerror(magma_int_t error)
{
    if (error > 0) return "function-specific error, see documentation";
    if (error < 0 && error > MAGMA_ERR) return "invalid argument";
    switch (error) {
    case MAGMA_SUCCESS:              return "success";
    case MAGMA_ERR:                  return "unknown error";
    case MAGMA_ERR_NOT_INITIALIZED:  return "not initialized";
    case MAGMA_ERR_NOT_SUPPORTED:    return "not supported";
    case MAGMA_ERR_HOST_ALLOC:       return "cannot allocate memory on CPU host";
    case MAGMA_ERR_DEVICE_ALLOC:     return "cannot allocate memory on GPU device";
    case MAGMA_ERR_INVALID_PTR:      return "invalid pointer";
    case MAGMA_ERR_UNKNOWN:          return "unspecified error";
    case MAGMA_ERR_NOT_IMPLEMENTED:  return "not implemented";
    case MAGMA_ERR_NAN:              return "NaN detected";
    default:                         return "unknown MAGMA error code";
    }
}

void magma_getdevice(magma_device_t *dev)
{
    int d = 0;
    hipGetDevice(&d);
    *dev = d;
}

void magma_setdevice(magma_device_t dev) { hipSetDevice(dev); }

void magma_device_sync(void) { hipDeviceSynchronize(); }

magma_int_t magma_malloc(magma_ptr *ptr_ptr, size_t bytes)
{
    if (bytes == 0) bytes = 8;   // like the reference: a valid pointer for size 0
    if (hipMalloc(ptr_ptr, bytes) != hipSuccess) {
        *ptr_ptr = nullptr;
        return MAGMA_ERR_DEVICE_ALLOC;
    }
    return MAGMA_SUCCESS;
}

magma_int_t magma_malloc_cpu(void **ptr_ptr, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    if (posix_memalign(ptr_ptr, 64, bytes) != 0) {
        *ptr_ptr = nullptr;
        return MAGMA_ERR_HOST_ALLOC;
    }
    return MAGMA_SUCCESS;
}

magma_int_t magma_malloc_pinned(void **ptr_ptr, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    if (hipHostMalloc(ptr_ptr, bytes, hipHostMallocDefault) != hipSuccess) {
        *ptr_ptr = nullptr;
        return MAGMA_ERR_HOST_ALLOC;
    }
    return MAGMA_SUCCESS;
}

magma_int_t magma_free_cpu(void *ptr)
{
    std::free(ptr);
    return MAGMA_SUCCESS;
}

magma_int_t magma_free_internal(magma_ptr ptr, const char *func, const char *file, int line)
{
    if (hipFree(ptr) != hipSuccess) {
        std::fprintf(stderr, "magma_free: invalid pointer %p in %s (%s:%d)\n", ptr, func, file, line);
        return MAGMA_ERR_INVALID_PTR;
    }
    return MAGMA_SUCCESS;
}

magma_int_t magma_free_pinned_internal(void *ptr, const char *func, const char *file, int line)
{
    if (hipHostFree(ptr) != hipSuccess) {
        std::fprintf(stderr, "magma_free_pinned: invalid pointer %p in %s (%s:%d)\n", ptr, func, file, line);
        return MAGMA_ERR_INVALID_PTR;
    }
    return MAGMA_SUCCESS;
}

void magma_queue_create_internal(magma_device_t device, magma_queue_t *queue_ptr,
                                 const char *func, const char *file, int line)
{
    magma_queue_t q = new magma_queue();
    q->device = device;
    hipSetDevice(device);
    if (hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking) != hipSuccess) {
        std::fprintf(stderr, "magma_queue_create: hipStreamCreate failed in %s (%s:%d)\n", func, file, line);
        delete q;
        *queue_ptr = nullptr;
        return;
    }
    q->own_stream = true;
    *queue_ptr = q;
}

void magma_queue_create_from_hip_internal(magma_device_t device, magma_hip_stream_t stream,
                                          magma_queue_t *queue_ptr,
                                          const char *func, const char *file, int line)
{
    (void)func; (void)file; (void)line;
    magma_queue_t q = new magma_queue();
    q->device = device;
    q->stream = reinterpret_cast<hipStream_t>(stream);
    q->own_stream = false;
    *queue_ptr = q;
}

magma_hip_stream_t magma_queue_get_hip_stream(magma_queue_t queue)
{
    return reinterpret_cast<magma_hip_stream_t>(queue->stream);
}

void magma_queue_destroy_internal(magma_queue_t queue, const char *func, const char *file, int line)
{
    (void)func; (void)file; (void)line;
    if (!queue) return;
    (void)hipStreamSynchronize(queue->stream);
    if (queue->work) hipFree(queue->work);
    if (queue->host) hipHostFree(queue->host);
    if (queue->own_stream) hipStreamDestroy(queue->stream);
    delete queue;
}

void magma_queue_sync_internal(magma_queue_t queue, const char *func, const char *file, int line)
{
    hipError_t e = hipStreamSynchronize(queue ? queue->stream : nullptr);
    if (e != hipSuccess)
        std::fprintf(stderr, "magma_queue_sync: %s in %s (%s:%d)\n", hipGetErrorString(e), func, file, line);
}

magma_int_t magma_queue_get_device(magma_queue_t queue) { return queue->device; }

void magma_queue_set_async(magma_queue_t queue, magma_int_t async) { queue->async = (async != 0); }

void magma_event_create(magma_event_t *event_ptr)
{
    hipEvent_t e;
    hipEventCreate(&e);
    *event_ptr = reinterpret_cast<magma_event_t>(e);
}

void magma_event_destroy(magma_event_t event) { hipEventDestroy(reinterpret_cast<hipEvent_t>(event)); }

void magma_event_record(magma_event_t event, magma_queue_t queue)
{
    hipEventRecord(reinterpret_cast<hipEvent_t>(event), queue->stream);
}

void magma_event_sync(magma_event_t event) { hipEventSynchronize(reinterpret_cast<hipEvent_t>(event)); }

float magma_event_elapsed_ms(magma_event_t start, magma_event_t stop)
{
    float ms = 0.f;
    hipEventElapsedTime(&ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(stop));
    return ms;
}

void magma_setvector_internal(magma_int_t n, magma_int_t elemSize, const void *hx_src, magma_int_t incx,
                              magma_ptr dy_dst, magma_int_t incy, magma_queue_t queue,
                              const char *func, const char *file, int line)
{
    hipError_t e = hipMemcpy2DAsync(dy_dst, (size_t)incy * elemSize, hx_src, (size_t)incx * elemSize,
                                    elemSize, n, hipMemcpyHostToDevice, stream_of(queue));
    if (e == hipSuccess) e = hipStreamSynchronize(stream_of(queue));
    if (e != hipSuccess) report_hip_error(e, func, file, line);
}

void magma_getvector_internal(magma_int_t n, magma_int_t elemSize, magma_const_ptr dx_src, magma_int_t incx,
                              void *hy_dst, magma_int_t incy, magma_queue_t queue,
                              const char *func, const char *file, int line)
{
    hipError_t e = hipMemcpy2DAsync(hy_dst, (size_t)incy * elemSize, dx_src, (size_t)incx * elemSize,
                                    elemSize, n, hipMemcpyDeviceToHost, stream_of(queue));
    if (e == hipSuccess) e = hipStreamSynchronize(stream_of(queue));
    if (e != hipSuccess) report_hip_error(e, func, file, line);
}

void magma_setmatrix_internal(magma_int_t m, magma_int_t n, magma_int_t elemSize,
                              const void *hA_src, magma_int_t lda, magma_ptr dB_dst, magma_int_t lddb,
                              magma_queue_t queue, const char *func, const char *file, int line)
{
    hipError_t e = hipMemcpy2DAsync(dB_dst, (size_t)lddb * elemSize, hA_src, (size_t)lda * elemSize,
                                    (size_t)m * elemSize, n, hipMemcpyHostToDevice, stream_of(queue));
    if (e == hipSuccess) e = hipStreamSynchronize(stream_of(queue));
    if (e != hipSuccess) report_hip_error(e, func, file, line);
}

void magma_getmatrix_internal(magma_int_t m, magma_int_t n, magma_int_t elemSize,
                              magma_const_ptr dA_src, magma_int_t ldda, void *hB_dst, magma_int_t ldb,
                              magma_queue_t queue, const char *func, const char *file, int line)
{
    hipError_t e = hipMemcpy2DAsync(hB_dst, (size_t)ldb * elemSize, dA_src, (size_t)ldda * elemSize,
                                    (size_t)m * elemSize, n, hipMemcpyDeviceToHost, stream_of(queue));
    if (e == hipSuccess) e = hipStreamSynchronize(stream_of(queue));
    if (e != hipSuccess) report_hip_error(e, func, file, line);
}

}  // extern "C"

// ---- pointer-array helpers (magmablas/zset_pointer.cu, magmablas/set_pointer.cu) ----
template <typename T>
__global__ void set_pointer_kernel(T **out, T *base, int64_t lda, int64_t row, int64_t col,
                                   int64_t batch_offset, int count)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = base + row + col * lda + (int64_t)i * batch_offset;
}

template <typename T>
__global__ void displace_pointer_kernel(T **out, T **in, int64_t lda, int64_t row, int64_t col, int count)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = in[i] + row + col * lda;
}

extern "C" {

void magma_dset_pointer(double **output_array, double *input, magma_int_t lda,
                        magma_int_t row, magma_int_t column, magma_int_t batch_offset,
                        magma_int_t batchCount, magma_queue_t queue)
{
    if (batchCount <= 0) return;
    set_pointer_kernel<double><<<ceil_div(batchCount, 256), 256, 0, stream_of(queue)>>>(
        output_array, input, lda, row, column, batch_offset, batchCount);
}

void magma_ddisplace_pointers(double **output_array, double **input_array, magma_int_t lda,
                              magma_int_t row, magma_int_t column,
                              magma_int_t batchCount, magma_queue_t queue)
{
    if (batchCount <= 0) return;
    displace_pointer_kernel<double><<<ceil_div(batchCount, 256), 256, 0, stream_of(queue)>>>(
        output_array, input_array, lda, row, column, batchCount);
}

void magma_iset_pointer(magma_int_t **output_array, magma_int_t *input, magma_int_t lda,
                        magma_int_t row, magma_int_t column, magma_int_t batchSize,
                        magma_int_t batchCount, magma_queue_t queue)
{
    if (batchCount <= 0) return;
    set_pointer_kernel<magma_int_t><<<ceil_div(batchCount, 256), 256, 0, stream_of(queue)>>>(
        output_array, input, lda, row, column, batchSize, batchCount);
}

}  // extern "C"
