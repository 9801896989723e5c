// internal.h -- shared definitions of libmagma_amd (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>

#include "magma_v2.h"

// Queue: one HIP stream on one device plus a growable device workspace and a
// small pinned host buffer for size readbacks.  The reference's queue carries a
// cuBLAS and a cuSPARSE handle as well (control/magma_internal.h:86-137); this
// library calls no vendor BLAS, so it carries none.
struct magma_queue {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    void *work = nullptr;          // device workspace, reused across calls
    size_t work_bytes = 0;
    void *host = nullptr;          // pinned host staging
    size_t host_bytes = 0;
    bool async = false;            // skip the end-of-routine synchronization
};

namespace magma_amd {

// Returns a device workspace of at least `bytes`, growing it when needed.
// Growth synchronizes the queue first, so work still in flight never sees its
// buffer freed.  Returns nullptr on allocation failure.
void *queue_workspace(magma_queue_t queue, size_t bytes);
// Pinned host buffer of at least `bytes` (same growth rule).
void *queue_host(magma_queue_t queue, size_t bytes);

inline hipStream_t stream_of(magma_queue_t q) { return q ? q->stream : nullptr; }

// Reports a failed HIP call (launch or runtime) on stderr; returns MAGMA_ERR_UNKNOWN.
magma_int_t report_hip_error(hipError_t e, const char *what, const char *file, int line);

inline size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

}  // namespace magma_amd

#define MAGMA_HIP_CHECK(expr)                                                               \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) return magma_amd::report_hip_error(e_, #expr, __FILE__, __LINE__); \
    } while (0)

#define MAGMA_LAUNCH_CHECK()                                                                \
    do {                                                                                    \
        hipError_t e_ = hipGetLastError();                                                  \
        if (e_ != hipSuccess) return magma_amd::report_hip_error(e_, "kernel launch", __FILE__, __LINE__); \
    } while (0)

namespace magma_amd {

// The reference drivers end with magma_queue_sync (src/zgetrf_batched.cpp:354):
// results are ready when the call returns, unless the queue was made
// asynchronous for graph capture.
inline magma_int_t sync_unless_async(magma_queue_t queue)
{
    if (!(queue && queue->async)) MAGMA_HIP_CHECK(hipStreamSynchronize(stream_of(queue)));
    return MAGMA_SUCCESS;
}

}  // namespace magma_amd
