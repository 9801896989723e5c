// workqueue.h -- device-side size buckets and tile lists for vbatched calls.
#pragma once

#include "internal.h"

namespace magma_amd {

// Size classes by max(m, n): <=8, <=16, <=32, <=64, <=128, <=256, <=512,
// <=1024, <=2048, <=8192 (the largest order the factor kernels take; callers
// reject larger ones before binning).
constexpr int kNumBuckets = 10;
constexpr int kBucketLimit[kNumBuckets] = {8, 16, 32, 64, 128, 256, 512, 1024, 2048, 8192};

// Views into one device allocation of workqueue_bytes(count):
//   ids[count]            matrix ids sorted by size class
//   counts[kNumBuckets]   matrices per class
//   offsets[kNumBuckets+1] start of each class in ids (exclusive prefix)
//   cursor[kNumBuckets]   scatter cursors (scratch)
struct WorkQueue {
    int *ids = nullptr;
    int *counts = nullptr;
    int *offsets = nullptr;
    int *cursor = nullptr;
};

size_t workqueue_bytes(int count);
WorkQueue workqueue_view(void *mem, int count);
// m may be null (square: m = n).  Enqueued on st; no host synchronization.
magma_int_t build_workqueue(const int *m, const int *n, int count, WorkQueue q, hipStream_t st);
// The same for the square matrices named by list[0 .. *count), where list and
// count live on the device; ids must have room for the longest such list.
magma_int_t build_workqueue_list(const int *n, const int *list, const int *count, WorkQueue q, hipStream_t st);
// prefix[count+1] of ceil(m_i/tile_m)*ceil(n_i/tile_n); enqueued on st.
magma_int_t build_tile_prefix(const int *m, const int *n, int count, int tile_m, int tile_n, int *prefix,
                              hipStream_t st);

}  // namespace magma_amd
