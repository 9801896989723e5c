/*
 * magma_types.h -- scalar, enum, error-code and queue types of the MAGMA C API
 * subset this library exports.  Values are those of the reference
 * (include/magma_types.h:44-47 magma_int_t, :333-349 error codes,
 * :397-419 enums) so that code written against MAGMA compiles and links
 * unchanged.  No HIP or CUDA header is needed to include this file.
 */
#ifndef MAGMA_AMD_TYPES_H
#define MAGMA_AMD_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/magma_types.h:44-47: 32-bit unless built for ILP64 (not supported here) */
typedef int          magma_int_t;
typedef int          magma_index_t;
typedef unsigned int magma_uindex_t;
typedef double       real_Double_t;
typedef magma_int_t  magma_device_t;

/* Opaque queue: a HIP stream on one device plus per-queue workspace.
 * (reference: struct magma_queue, control/magma_internal.h:86-137) */
struct magma_queue;
typedef struct magma_queue *magma_queue_t;

/* Layout-compatible with hipStream_t / hipEvent_t without including HIP. */
typedef struct ihipStream_t *magma_hip_stream_t;
typedef struct ihipEvent_t  *magma_event_t;

typedef void         *magma_ptr;
typedef const void   *magma_const_ptr;
typedef double       *magmaDouble_ptr;
typedef const double *magmaDouble_const_ptr;
typedef float        *magmaFloat_ptr;
typedef magma_int_t  *magmaInt_ptr;

/* ---- return codes (include/magma_types.h:333-349) -------------------- */
#define MAGMA_SUCCESS               0
#define MAGMA_ERR                  -100
#define MAGMA_ERR_NOT_INITIALIZED  -101
#define MAGMA_ERR_REINITIALIZED    -102
#define MAGMA_ERR_NOT_SUPPORTED    -103
#define MAGMA_ERR_ILLEGAL_VALUE    -104
#define MAGMA_ERR_NOT_FOUND        -105
#define MAGMA_ERR_ALLOCATION       -106
#define MAGMA_ERR_INTERNAL_LIMIT   -107
#define MAGMA_ERR_UNALLOCATED      -108
#define MAGMA_ERR_FILESYSTEM       -109
#define MAGMA_ERR_UNEXPECTED       -110
#define MAGMA_ERR_SEQUENCE_FLUSHED -111
#define MAGMA_ERR_HOST_ALLOC       -112
#define MAGMA_ERR_DEVICE_ALLOC     -113
#define MAGMA_ERR_CUDASTREAM       -114
#define MAGMA_ERR_INVALID_PTR      -115
#define MAGMA_ERR_UNKNOWN          -116
#define MAGMA_ERR_NOT_IMPLEMENTED  -117
#define MAGMA_ERR_NAN              -118

/* ---- parameter constants (include/magma_types.h:397-419) -------------- */
typedef enum { MagmaFalse = 0, MagmaTrue = 1 } magma_bool_t;
typedef enum { MagmaRowMajor = 101, MagmaColMajor = 102 } magma_order_t;
typedef enum {
    MagmaNoTrans = 111, MagmaTrans = 112, MagmaConjTrans = 113,
    Magma_ConjTrans = MagmaConjTrans
} magma_trans_t;
typedef enum {
    MagmaUpper = 121, MagmaLower = 122, MagmaFull = 123, MagmaHessenberg = 124
} magma_uplo_t;
typedef enum { MagmaNonUnit = 131, MagmaUnit = 132 } magma_diag_t;
typedef enum { MagmaLeft = 141, MagmaRight = 142, MagmaBothSides = 143 } magma_side_t;

#ifdef __cplusplus
}
#endif
#endif
