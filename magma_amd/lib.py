"""ctypes binding of libmagma_amd.so (the C-ABI library built from magma_amd/csrc).

The library is the product; this module only declares its exported C
functions (prototypes in include/*.h) so that Python tests and benchmarks can
call them exactly as a C caller would.  Loading fails loudly when the library
has not been built -- there is no fallback implementation.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmagma_amd.so")

c_int = ctypes.c_int
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
c_char_p = ctypes.c_char_p
c_float = ctypes.c_float

# enum values (include/magma_types.h)
MagmaNoTrans, MagmaTrans, MagmaConjTrans = 111, 112, 113
MagmaUpper, MagmaLower, MagmaFull = 121, 122, 123
MagmaNonUnit, MagmaUnit = 131, 132
MagmaLeft, MagmaRight = 141, 142
MAGMA_SUCCESS = 0
MAGMA_ERR_DEVICE_ALLOC = -113
MAGMA_ERR_NOT_SUPPORTED = -103
MAGMA_ERR_NOT_IMPLEMENTED = -117
MAGMA_MIXED_ITERMAX = 10
MAGMA_AMD_SOLVE_LU, MAGMA_AMD_SOLVE_SPD = 0, 1
MAGMA_AMD_TUNE_GETRF, MAGMA_AMD_TUNE_MIXED, MAGMA_AMD_TUNE_GEMM, MAGMA_AMD_TUNE_VERBOSE = 1, 2, 4, 256


class Request(ctypes.Structure):
    """magma_amd_request_t (include/magma_amd.h)"""
    _fields_ = [("kind", c_int), ("n", c_int), ("nrhs", c_int),
                ("dA", ctypes.c_void_p), ("ldda", c_int),
                ("dB", ctypes.c_void_p), ("lddb", c_int),
                ("dX", ctypes.c_void_p), ("lddx", c_int),
                ("factor_key", ctypes.c_longlong)]


class Completion(ctypes.Structure):
    """magma_amd_completion_t"""
    _fields_ = [("ticket", ctypes.c_longlong), ("info", c_int), ("iter", c_int), ("reused", c_int),
                ("berr", c_double), ("t_submit", c_double), ("t_start", c_double), ("t_done", c_double)]


class ServerStats(ctypes.Structure):
    """magma_amd_server_stats_t"""
    _fields_ = [("submitted", ctypes.c_longlong), ("completed", ctypes.c_longlong),
                ("batches", ctypes.c_longlong), ("evictions", ctypes.c_longlong),
                ("budget_bytes", ctypes.c_size_t), ("bytes_in_use", ctypes.c_size_t),
                ("bytes_peak", ctypes.c_size_t), ("resident", c_int), ("max_batch_seen", c_int)]

# name -> (restype, argtypes); every pointer argument is passed as an integer address
_P = c_void_p
PROTOTYPES = {
    # runtime
    "magma_init": (c_int, []),
    "magma_finalize": (c_int, []),
    "magma_version": (None, [_P, _P, _P]),
    "magma_wtime": (c_double, []),
    "magma_sync_wtime": (c_double, [_P]),
    "magma_xerbla": (None, [c_char_p, c_int]),
    "magma_strerror": (c_char_p, [c_int]),
    "magma_getdevice": (None, [_P]),
    "magma_setdevice": (None, [c_int]),
    "magma_device_sync": (None, []),
    "magma_malloc": (c_int, [_P, ctypes.c_size_t]),
    "magma_malloc_cpu": (c_int, [_P, ctypes.c_size_t]),
    "magma_malloc_pinned": (c_int, [_P, ctypes.c_size_t]),
    "magma_free_cpu": (c_int, [_P]),
    "magma_free_internal": (c_int, [_P, c_char_p, c_char_p, c_int]),
    "magma_free_pinned_internal": (c_int, [_P, c_char_p, c_char_p, c_int]),
    "magma_queue_create_internal": (None, [c_int, _P, c_char_p, c_char_p, c_int]),
    "magma_queue_create_from_hip_internal": (None, [c_int, _P, _P, c_char_p, c_char_p, c_int]),
    "magma_queue_destroy_internal": (None, [_P, c_char_p, c_char_p, c_int]),
    "magma_queue_sync_internal": (None, [_P, c_char_p, c_char_p, c_int]),
    "magma_queue_get_device": (c_int, [_P]),
    "magma_queue_get_hip_stream": (_P, [_P]),
    "magma_queue_set_async": (None, [_P, c_int]),
    "magma_event_create": (None, [_P]),
    "magma_event_destroy": (None, [_P]),
    "magma_event_record": (None, [_P, _P]),
    "magma_event_sync": (None, [_P]),
    "magma_event_elapsed_ms": (c_float, [_P, _P]),
    "magma_setvector_internal": (None, [c_int, c_int, _P, c_int, _P, c_int, _P, c_char_p, c_char_p, c_int]),
    "magma_getvector_internal": (None, [c_int, c_int, _P, c_int, _P, c_int, _P, c_char_p, c_char_p, c_int]),
    "magma_setmatrix_internal": (None, [c_int, c_int, c_int, _P, c_int, _P, c_int, _P, c_char_p, c_char_p, c_int]),
    "magma_getmatrix_internal": (None, [c_int, c_int, c_int, _P, c_int, _P, c_int, _P, c_char_p, c_char_p, c_int]),
    "magma_dset_pointer": (None, [_P, _P, c_int, c_int, c_int, c_int, c_int, _P]),
    "magma_ddisplace_pointers": (None, [_P, _P, c_int, c_int, c_int, c_int, _P]),
    "magma_iset_pointer": (None, [_P, _P, c_int, c_int, c_int, c_int, c_int, _P]),
    # batched factor / solve (include/magma_dbatched.h)
    "magma_dgetrf_batched": (c_int, [c_int, c_int, _P, c_int, _P, _P, c_int, _P]),
    "magma_dgetrs_batched": (c_int, [c_int, c_int, c_int, _P, c_int, _P, _P, c_int, c_int, _P]),
    "magma_dgesv_batched": (c_int, [c_int, c_int, _P, c_int, _P, _P, c_int, _P, c_int, _P]),
    "magma_dpotrf_batched": (c_int, [c_int, c_int, _P, c_int, _P, c_int, _P]),
    "magma_dpotrs_batched": (c_int, [c_int, c_int, c_int, _P, c_int, _P, c_int, c_int, _P]),
    "magma_dposv_batched": (c_int, [c_int, c_int, c_int, _P, c_int, _P, c_int, _P, c_int, _P]),
    "magma_dgemm_batched": (None, [c_int, c_int, c_int, c_int, c_int, c_double, _P, c_int, _P, c_int,
                                   c_double, _P, c_int, c_int, _P]),
    "magmablas_dgemm_batched": (None, [c_int, c_int, c_int, c_int, c_int, c_double, _P, c_int, _P, c_int,
                                       c_double, _P, c_int, c_int, _P]),
    "magma_dgetrf_nopiv_batched": (c_int, [c_int, c_int, _P, c_int, _P, c_int, _P]),
    "magma_dgetrs_nopiv_batched": (c_int, [c_int, c_int, c_int, _P, c_int, _P, c_int, _P, c_int, _P]),
    "magma_dgesv_nopiv_batched": (c_int, [c_int, c_int, _P, c_int, _P, c_int, _P, c_int, _P]),
    "magma_dgeqrf_batched": (c_int, [c_int, c_int, _P, c_int, _P, _P, c_int, _P]),
    "magma_dgesv_rbt_batched": (c_int, [c_int, c_int, _P, c_int, _P, c_int, _P, c_int, _P]),
    "magma_dgetri_outofplace_batched": (c_int, [c_int, _P, c_int, _P, _P, c_int, _P, c_int, _P]),
    "magmablas_dtrsm_batched": (None, [c_int, c_int, c_int, c_int, c_int, c_int, c_double, _P, c_int, _P, c_int,
                                       c_int, _P]),
    "magmablas_dtrsv_batched": (None, [c_int, c_int, c_int, c_int, _P, c_int, _P, c_int, c_int, _P]),
    "magmablas_dgemv_batched": (None, [c_int, c_int, c_int, c_double, _P, c_int, _P, c_int, c_double, _P, c_int,
                                       c_int, _P]),
    # mixed precision (include/magma_amd.h)
    "magma_dsgesv_batched": (c_int, [c_int, c_int, _P, c_int, _P, _P, c_int, _P, c_int, _P, _P, c_int, _P]),
    "magma_dsposv_batched": (c_int, [c_int, c_int, c_int, _P, c_int, _P, c_int, _P, c_int, _P, _P, c_int, _P]),
    "magma_dsgesv_vbatched": (c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, c_int, _P]),
    "magma_dsposv_vbatched": (c_int, [c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, c_int, _P]),
    # vbatched (include/magma_dvbatched.h)
    "magma_dgetrf_vbatched": (c_int, [_P, _P, _P, _P, _P, _P, c_int, _P]),
    "magma_dgetrf_vbatched_max_nocheck": (c_int, [_P, _P, c_int, _P, _P, _P, _P, c_int, _P]),
    "magma_dgetrf_vbatched_max_nocheck_noalloc": (c_int, [_P, _P, c_int, _P, _P, _P, _P, c_int, _P]),
    "magma_dgetrs_vbatched": (c_int, [c_int, _P, _P, _P, _P, _P, _P, _P, c_int, _P]),
    "magma_dpotrf_vbatched": (c_int, [c_int, _P, _P, _P, _P, c_int, _P]),
    "magma_dpotrf_vbatched_max_nocheck": (c_int, [c_int, _P, _P, _P, _P, c_int, c_int, _P]),
    "magma_dpotrs_vbatched": (c_int, [c_int, _P, _P, _P, _P, _P, _P, c_int, _P]),
    "magma_dgetrs_vbatched_max_nocheck": (c_int, [c_int, _P, _P, _P, _P, _P, _P, _P, c_int, c_int, _P]),
    "magma_dpotrs_vbatched_max_nocheck": (c_int, [c_int, _P, _P, _P, _P, _P, _P, c_int, c_int, _P]),
    "magmablas_dgemm_vbatched": (None, [c_int, c_int, _P, _P, _P, c_double, _P, _P, _P, _P, c_double,
                                        _P, _P, c_int, _P]),
    "magmablas_dgemm_vbatched_max_nocheck": (None, [c_int, c_int, _P, _P, _P, c_double, _P, _P, _P, _P,
                                                    c_double, _P, _P, c_int, c_int, c_int, c_int, _P]),
    "magmablas_dgemm_vbatched_max": (None, [c_int, c_int, _P, _P, _P, c_double, _P, _P, _P, _P, c_double,
                                            _P, _P, c_int, c_int, c_int, c_int, _P]),
    "magmablas_dgemm_vbatched_nocheck": (None, [c_int, c_int, _P, _P, _P, c_double, _P, _P, _P, _P,
                                                c_double, _P, _P, c_int, _P]),
    "magma_amd_autotune": (c_int, [_P, c_int]),
    "magma_amd_tune_get": (c_int, [c_int, c_int]),
    "magma_amd_tune_set": (c_int, [c_int, c_int, c_int]),
    "magma_amd_tune_save": (c_int, [c_char_p]),
    "magma_amd_tune_load": (c_int, [c_char_p]),
    "magma_amd_server_create": (c_int, [c_int, ctypes.c_size_t, c_int, _P]),
    "magma_amd_server_submit": (ctypes.c_longlong, [_P, _P]),
    "magma_amd_server_poll": (c_int, [_P, _P, c_int]),
    "magma_amd_server_wait": (c_int, [_P, _P, c_int, c_double]),
    "magma_amd_server_drain": (c_int, [_P]),
    "magma_amd_server_stats": (c_int, [_P, _P]),
    "magma_amd_server_run_trace": (c_int, [_P, _P, _P, c_int, _P]),
    "magma_amd_server_destroy": (None, [_P]),
}

OPTIONAL = set()


class MagmaLibrary:
    """Loaded libmagma_amd.so with typed entry points as attributes."""

    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise RuntimeError(
                f"libmagma_amd.so not found at {path}; build it first "
                "(python -c 'import __graft_entry__ as g; g.build()' or make -C magma_amd/csrc)")
        self.path = path
        self.cdll = ctypes.CDLL(path)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(self.cdll, name, None)
            if fn is None:
                if name in OPTIONAL:
                    continue
                raise RuntimeError(f"{path} does not export {name}")
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    def exported(self, name):
        return hasattr(self.cdll, name)


_lib = None


def load():
    """Return the process-wide MagmaLibrary (loads it on first use).

    torch, when installed, is imported first: torch ships its own HIP runtime
    under the same soname as the system one, and a process must use a single
    runtime so that torch's streams and allocations are valid in the library."""
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = MagmaLibrary()
    return _lib
