/*
 * magma_v2.h -- umbrella header of libmagma_amd: runtime (init, queues,
 * memory, copies, errors) plus the batched / vbatched double-precision
 * factor-and-solve API.  Prototypes mirror the reference's
 * include/magma_auxiliary.h and include/magma_copy.h (cited per function);
 * the queue wraps a HIP stream instead of a CUDA stream + cuBLAS handle.
 */
#ifndef MAGMA_AMD_V2_H
#define MAGMA_AMD_V2_H

#include "magma_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- initialization (include/magma_auxiliary.h:23-24; interface_cuda/interface.cpp:168,246) */
magma_int_t magma_init(void);
magma_int_t magma_finalize(void);
void magma_version(magma_int_t *major, magma_int_t *minor, magma_int_t *micro);

/* ---- timing (include/magma_auxiliary.h:47-48) */
real_Double_t magma_wtime(void);
real_Double_t magma_sync_wtime(magma_queue_t queue);

/* ---- errors (control/xerbla.cpp:51; interface_cuda/error.cpp) */
void        magma_xerbla(const char *srname, magma_int_t neg_info);
const char *magma_strerror(magma_int_t error);

/* ---- devices (include/magma_auxiliary.h:220-226) */
void magma_getdevice(magma_device_t *dev);
void magma_setdevice(magma_device_t dev);
void magma_device_sync(void);

/* ---- memory (include/magma_auxiliary.h:83-110) */
magma_int_t magma_malloc(magma_ptr *ptr_ptr, size_t bytes);
magma_int_t magma_malloc_cpu(void **ptr_ptr, size_t bytes);
magma_int_t magma_malloc_pinned(void **ptr_ptr, size_t bytes);
magma_int_t magma_free_cpu(void *ptr);
magma_int_t magma_free_internal(magma_ptr ptr, const char *func, const char *file, int line);
magma_int_t magma_free_pinned_internal(void *ptr, const char *func, const char *file, int line);
#define magma_free(ptr)        magma_free_internal(ptr, __func__, __FILE__, __LINE__)
#define magma_free_pinned(ptr) magma_free_pinned_internal(ptr, __func__, __FILE__, __LINE__)

static inline magma_int_t magma_imalloc(magma_int_t **p, size_t n) { return magma_malloc((magma_ptr *)p, n * sizeof(magma_int_t)); }
static inline magma_int_t magma_dmalloc(double **p, size_t n)      { return magma_malloc((magma_ptr *)p, n * sizeof(double)); }
static inline magma_int_t magma_smalloc(float **p, size_t n)       { return magma_malloc((magma_ptr *)p, n * sizeof(float)); }

/* ---- queues (include/magma_auxiliary.h:232-280; interface_cuda/interface.cpp:690-871) */
void magma_queue_create_internal(magma_device_t device, magma_queue_t *queue_ptr,
                                 const char *func, const char *file, int line);
void magma_queue_destroy_internal(magma_queue_t queue, const char *func, const char *file, int line);
void magma_queue_sync_internal(magma_queue_t queue, const char *func, const char *file, int line);
magma_int_t magma_queue_get_device(magma_queue_t queue);
#define magma_queue_create(device, queue_ptr) \
        magma_queue_create_internal(device, queue_ptr, __func__, __FILE__, __LINE__)
#define magma_queue_destroy(queue) \
        magma_queue_destroy_internal(queue, __func__, __FILE__, __LINE__)
#define magma_queue_sync(queue) \
        magma_queue_sync_internal(queue, __func__, __FILE__, __LINE__)

/* Extension (counterpart of magma_queue_create_from_cuda_internal,
 * include/magma_auxiliary.h:250-258): wrap an existing HIP stream; the queue
 * does not own or destroy it. */
void magma_queue_create_from_hip_internal(magma_device_t device, magma_hip_stream_t stream,
                                          magma_queue_t *queue_ptr,
                                          const char *func, const char *file, int line);
#define magma_queue_create_from_hip(device, stream, queue_ptr) \
        magma_queue_create_from_hip_internal(device, stream, queue_ptr, __func__, __FILE__, __LINE__)
magma_hip_stream_t magma_queue_get_hip_stream(magma_queue_t queue);
/* Extension: with async != 0 the factor/solve routines return as soon as their
 * work is enqueued instead of synchronizing the queue (the reference's
 * behaviour, kept as the default).  Needed for HIP graph capture and for
 * back-to-back submission; the caller then synchronizes. */
void magma_queue_set_async(magma_queue_t queue, magma_int_t async);

/* ---- events (include/magma_auxiliary.h:284-300) */
void magma_event_create(magma_event_t *event_ptr);
void magma_event_destroy(magma_event_t event);
void magma_event_record(magma_event_t event, magma_queue_t queue);
void magma_event_sync(magma_event_t event);
/* Extension: milliseconds between two recorded events (hipEventElapsedTime). */
float magma_event_elapsed_ms(magma_event_t start, magma_event_t stop);

/* ---- copies (include/magma_copy.h:45-130) */
void magma_setvector_internal(magma_int_t n, magma_int_t elemSize, const void *hx_src, magma_int_t incx,
                              magma_ptr dy_dst, magma_int_t incy, magma_queue_t queue,
                              const char *func, const char *file, int line);
void magma_getvector_internal(magma_int_t n, magma_int_t elemSize, magma_const_ptr dx_src, magma_int_t incx,
                              void *hy_dst, magma_int_t incy, magma_queue_t queue,
                              const char *func, const char *file, int line);
void magma_setmatrix_internal(magma_int_t m, magma_int_t n, magma_int_t elemSize,
                              const void *hA_src, magma_int_t lda, magma_ptr dB_dst, magma_int_t lddb,
                              magma_queue_t queue, const char *func, const char *file, int line);
void magma_getmatrix_internal(magma_int_t m, magma_int_t n, magma_int_t elemSize,
                              magma_const_ptr dA_src, magma_int_t ldda, void *hB_dst, magma_int_t ldb,
                              magma_queue_t queue, const char *func, const char *file, int line);
#define magma_setvector(n, es, hx, incx, dy, incy, q) \
        magma_setvector_internal(n, es, hx, incx, dy, incy, q, __func__, __FILE__, __LINE__)
#define magma_getvector(n, es, dx, incx, hy, incy, q) \
        magma_getvector_internal(n, es, dx, incx, hy, incy, q, __func__, __FILE__, __LINE__)
#define magma_setmatrix(m, n, es, hA, lda, dB, lddb, q) \
        magma_setmatrix_internal(m, n, es, hA, lda, dB, lddb, q, __func__, __FILE__, __LINE__)
#define magma_getmatrix(m, n, es, dA, ldda, hB, ldb, q) \
        magma_getmatrix_internal(m, n, es, dA, ldda, hB, ldb, q, __func__, __FILE__, __LINE__)

#ifdef __cplusplus
}
#endif

#include "magma_dbatched.h"
#include "magma_dvbatched.h"
#include "magma_amd.h"

#endif
