"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
as the checker; the product never imports it.

Arrays follow the tests' storage convention: a batch of column-major matrices
is a C-contiguous numpy array of shape (count, cols, ld) (storage[b, j, i] =
A_b(i, j)); pivots are int32 (count, min(m, n)), 1-based.
"""
import ctypes
import glob
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_L = ctypes.c_int64
_C = ctypes.c_char

_PROTO = {
    "orc_idamax": (_I, [_I, _P]),
    "orc_dgetrf": (None, [_I, _I, _P, _I, _P, _P]),
    "orc_dlaswp": (None, [_I, _P, _I, _I, _I, _P, _I]),
    "orc_dgetrs": (None, [_C, _I, _I, _P, _I, _P, _P, _I]),
    "orc_dpotrf": (None, [_C, _I, _P, _I, _P]),
    "orc_dpotrs": (None, [_C, _I, _I, _P, _I, _P, _I]),
    "orc_dgemm": (None, [_C, _C, _I, _I, _I, _D, _P, _I, _P, _I, _D, _P, _I]),
    "orc_dgetrf_nopiv": (None, [_I, _I, _P, _I, _P]),
    "orc_dgeqrf": (None, [_I, _I, _P, _I, _P]),
    "orc_rbt_butterflies": (None, [_I, _P, _P]),
    "orc_dgesv_rbt": (None, [_I, _I, _P, _I, _P, _I, _P]),
    "orc_dtrsm": (None, [_C, _C, _C, _C, _I, _I, _D, _P, _I, _P, _I]),
    "orc_dtrsv": (None, [_C, _C, _C, _I, _P, _I, _P, _I]),
    "orc_dgemv": (None, [_C, _I, _I, _D, _P, _I, _P, _I, _D, _P, _I]),
    "orc_eps": (_D, []),
    "orc_dlange": (_D, [_C, _I, _I, _P, _I]),
    "orc_lu_error": (_D, [_I, _I, _P, _I, _P, _I, _P]),
    "orc_solve_error": (_D, [_I, _I, _P, _I, _P, _I, _P, _I]),
    "orc_chol_error": (_D, [_C, _I, _P, _I, _P, _I]),
    "orc_batch_dgetrf": (None, [_I, _P, _I, _L, _P, _P, _I, _I]),
    "orc_batch_dgesv": (None, [_I, _I, _P, _I, _L, _P, _P, _I, _L, _P, _I, _I]),
    "orc_batch_dposv": (None, [_C, _I, _I, _P, _I, _L, _P, _I, _L, _P, _I, _I]),
    "orc_vbatch_dgetrf": (None, [_P, _P, _P, _P, _P, _P, _P, _I, _I]),
    "orc_vbatch_dgemm": (None, [_P, _P, _P, _D, _P, _P, _P, _P, _P, _P, _D, _P, _P, _P, _I, _I]),
    "orc_lapack_open": (_I, [ctypes.c_char_p]),
    "orc_lapack_config": (ctypes.c_char_p, []),
    "orc_lapack_batch_dgesv": (_I, [_I, _I, _P, _I, _L, _P, _P, _I, _L, _P, _I, _I]),
    "orc_lapack_batch_dposv": (_I, [_C, _I, _I, _P, _I, _L, _P, _I, _L, _P, _I, _I]),
    "orc_lapack_batch_dgetrf": (_I, [_I, _P, _I, _L, _P, _P, _I, _I]),
    "orc_wtime": (_D, []),
    "orc_max_threads": (_I, []),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
        _lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _PROTO.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def _p(a):
    return a.ctypes.data_as(_P)


def _f64(a):
    a = np.require(a, dtype=np.float64, requirements=["C", "W"])
    return a


EPS = 2.0 ** -53


def dgetrf(A, m=None, n=None):
    """A: (cols, ld) column-major storage of one matrix (modified copy returned)."""
    A = _f64(np.array(A, dtype=np.float64, copy=True))
    cols, ld = A.shape
    m = ld if m is None else m
    n = cols if n is None else n
    ipiv = np.zeros(max(min(m, n), 1), dtype=np.int32)
    info = np.zeros(1, dtype=np.int32)
    lib().orc_dgetrf(m, n, _p(A), ld, _p(ipiv), _p(info))
    return A, ipiv[: min(m, n)], int(info[0])


def dgetrs(LU, ipiv, B, trans=b"N"):
    LU = _f64(np.array(LU, copy=True))
    B = _f64(np.array(B, copy=True))
    n = LU.shape[0]
    nrhs, ldb = B.shape
    ip = np.require(ipiv, dtype=np.int32, requirements=["C"])
    lib().orc_dgetrs(trans, n, nrhs, _p(LU), LU.shape[1], _p(ip), _p(B), ldb)
    return B


def dpotrf(A, uplo=b"L"):
    A = _f64(np.array(A, copy=True))
    n, ld = A.shape
    info = np.zeros(1, dtype=np.int32)
    lib().orc_dpotrf(uplo, n, _p(A), ld, _p(info))
    return A, int(info[0])


def dpotrs(L, B, uplo=b"L"):
    L = _f64(np.array(L, copy=True))
    B = _f64(np.array(B, copy=True))
    n, ld = L.shape
    nrhs, ldb = B.shape
    lib().orc_dpotrs(uplo, n, nrhs, _p(L), ld, _p(B), ldb)
    return B


def dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    A = _f64(np.array(A, copy=True)); B = _f64(np.array(B, copy=True)); C = _f64(np.array(C, copy=True))
    lib().orc_dgemm(ta, tb, m, n, k, alpha, _p(A), lda, _p(B), ldb, beta, _p(C), ldc)
    return C


def dgetrf_nopiv(A, m=None, n=None):
    """LU without pivoting of one (cols, ld) matrix; returns (LU, info)."""
    A = _f64(np.array(A, dtype=np.float64, copy=True))
    cols, ld = A.shape
    m = ld if m is None else m
    n = cols if n is None else n
    info = np.zeros(1, dtype=np.int32)
    lib().orc_dgetrf_nopiv(m, n, _p(A), ld, _p(info))
    return A, int(info[0])


def dgeqrf(A, m=None, n=None):
    """Householder QR of one (cols, ld) matrix; returns (QR, tau)."""
    A = _f64(np.array(A, dtype=np.float64, copy=True))
    cols, ld = A.shape
    m = ld if m is None else m
    n = cols if n is None else n
    tau = np.zeros(max(min(m, n), 1))
    lib().orc_dgeqrf(m, n, _p(A), ld, _p(tau))
    return A, tau[: min(m, n)]


def qr_errors(A, QR, tau, m, n):
    """The tester's two checks (testing/testing_zgeqrf_batched.cpp:34-82):
    ||R - Q'A||_1 / (n ||A||_1) and ||I - Q'Q||_1 / n, Q formed from the
    reflectors (numpy matrices; A and QR in (cols, ld) storage)."""
    Am = np.asarray(A).T[:m, :n]
    F = np.asarray(QR).T[:m, :n]
    k = min(m, n)
    Q = np.eye(m)
    for i in reversed(range(k)):
        v = np.zeros(m); v[i] = 1.0; v[i + 1:] = F[i + 1:, i]
        Q = Q - tau[i] * np.outer(v, v @ Q)
    Q = Q[:, :k]
    R = np.triu(F[:k, :])
    an = np.abs(Am).sum(axis=0).max() if Am.size else 0.0
    e1 = np.abs(R - Q.T @ Am).sum(axis=0).max() / (n * an) if n and an else 0.0
    e2 = np.abs(np.eye(k) - Q.T @ Q).sum(axis=0).max() / n if n else 0.0
    return e1, e2


def rbt_butterflies(n):
    u = np.zeros(max(2 * n, 1)); v = np.zeros(max(2 * n, 1))
    lib().orc_rbt_butterflies(n, _p(u), _p(v))
    return u[: 2 * n], v[: 2 * n]


def dgesv_rbt(A, B):
    """A: (n, lda) storage, B: (nrhs, ldb) storage; returns (A', X, info)."""
    A = _f64(np.array(A, copy=True)); B = _f64(np.array(B, copy=True))
    n, lda = A.shape
    nrhs, ldb = B.shape
    info = np.zeros(1, dtype=np.int32)
    lib().orc_dgesv_rbt(n, nrhs, _p(A), lda, _p(B), ldb, _p(info))
    return A, B, int(info[0])


def dtrsm(side, uplo, trans, diag, m, n, alpha, A, B):
    """A: (cols, lda) storage of the triangular matrix (order m for Left, n for
    Right); B: (n, ldb) storage of the m x n right-hand side.  Returns the
    solution in a copy of B."""
    A = _f64(np.array(A, copy=True)); B = _f64(np.array(B, copy=True))
    lib().orc_dtrsm(side, uplo, trans, diag, m, n, alpha, _p(A), A.shape[1], _p(B), B.shape[1])
    return B


def dtrsv(uplo, trans, diag, n, A, x, incx=1):
    """x: 1-D storage with increment incx; returns the solved copy."""
    A = _f64(np.array(A, copy=True)); x = _f64(np.array(x, dtype=np.float64, copy=True))
    lib().orc_dtrsv(uplo, trans, diag, n, _p(A), A.shape[1], _p(x), incx)
    return x


def dgemv(trans, m, n, alpha, A, x, beta, y, incx=1, incy=1):
    A = _f64(np.array(A, copy=True)); x = _f64(np.array(x, dtype=np.float64, copy=True))
    y = _f64(np.array(y, dtype=np.float64, copy=True))
    lib().orc_dgemv(trans, m, n, alpha, _p(A), A.shape[1], _p(x), incx, beta, _p(y), incy)
    return y


def idamax(x):
    x = _f64(np.array(x, dtype=np.float64))
    return lib().orc_idamax(len(x), _p(x))


def lu_error(A, LU, ipiv, m=None, n=None):
    A = _f64(np.array(A)); LU = _f64(np.array(LU))
    ip = np.require(ipiv, dtype=np.int32, requirements=["C"])
    cols, ld = A.shape
    m = ld if m is None else m
    n = cols if n is None else n
    return lib().orc_lu_error(m, n, _p(A), ld, _p(LU), LU.shape[1], _p(ip))


def solve_error(A, X, B):
    A = _f64(np.array(A)); X = _f64(np.array(X)); B = _f64(np.array(B))
    n, lda = A.shape
    nrhs = X.shape[0]
    return lib().orc_solve_error(n, nrhs, _p(A), lda, _p(X), X.shape[1], _p(B), B.shape[1])


def chol_error(A, L, uplo=b"L"):
    A = _f64(np.array(A)); L = _f64(np.array(L))
    n, lda = A.shape
    return lib().orc_chol_error(uplo, n, _p(A), lda, _p(L), L.shape[1])


def batch_dgetrf(A, nthreads=None):
    """A: (count, n, ld) column-major batch; returns (LU, ipiv, info)."""
    A = _f64(np.array(A, copy=True))
    count, n, ld = A.shape
    ipiv = np.zeros((count, n), dtype=np.int32)
    info = np.zeros(count, dtype=np.int32)
    lib().orc_batch_dgetrf(n, _p(A), ld, n * ld, _p(ipiv), _p(info), count, nthreads or os.cpu_count())
    return A, ipiv, info


def batch_dgesv(A, B, nthreads=None):
    A = _f64(np.array(A, copy=True)); B = _f64(np.array(B, copy=True))
    count, n, ld = A.shape
    _, nrhs, ldb = B.shape
    ipiv = np.zeros((count, n), dtype=np.int32)
    info = np.zeros(count, dtype=np.int32)
    lib().orc_batch_dgesv(n, nrhs, _p(A), ld, n * ld, _p(ipiv), _p(B), ldb, nrhs * ldb,
                          _p(info), count, nthreads or os.cpu_count())
    return A, ipiv, B, info


def batch_dposv(A, B, uplo=b"L", nthreads=None):
    A = _f64(np.array(A, copy=True)); B = _f64(np.array(B, copy=True))
    count, n, ld = A.shape
    _, nrhs, ldb = B.shape
    info = np.zeros(count, dtype=np.int32)
    lib().orc_batch_dposv(uplo, n, nrhs, _p(A), ld, n * ld, _p(B), ldb, nrhs * ldb,
                          _p(info), count, nthreads or os.cpu_count())
    return A, B, info


def find_openblas():
    """Path of an LP64 OpenBLAS exporting scipy_dgetrf_ (scipy's bundle), or None."""
    try:
        import scipy
        d = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
        c = sorted(glob.glob(os.path.join(d, "libscipy_openblas-*.so")))
        return c[0] if c else None
    except Exception:
        return None


_lapack_ok = None


def lapack():
    """Opens the third-party LAPACK once; returns its config string or None."""
    global _lapack_ok
    if _lapack_ok is None:
        path = find_openblas()
        _lapack_ok = bool(path) and lib().orc_lapack_open(path.encode()) == 0
    return lib().orc_lapack_config().decode() if _lapack_ok else None


def lapack_batch_dgesv(A, B, nthreads=None):
    if lapack() is None:
        raise RuntimeError("no third-party LAPACK found")
    A = _f64(np.array(A, copy=True)); B = _f64(np.array(B, copy=True))
    count, n, ld = A.shape
    _, nrhs, ldb = B.shape
    ipiv = np.zeros((count, n), dtype=np.int32)
    info = np.zeros(count, dtype=np.int32)
    lib().orc_lapack_batch_dgesv(n, nrhs, _p(A), ld, n * ld, _p(ipiv), _p(B), ldb, nrhs * ldb,
                                 _p(info), count, nthreads or os.cpu_count())
    return A, ipiv, B, info


def lapack_batch_dposv(A, B, uplo=b"L", nthreads=None):
    if lapack() is None:
        raise RuntimeError("no third-party LAPACK found")
    A = _f64(np.array(A, copy=True)); B = _f64(np.array(B, copy=True))
    count, n, ld = A.shape
    _, nrhs, ldb = B.shape
    info = np.zeros(count, dtype=np.int32)
    lib().orc_lapack_batch_dposv(uplo, n, nrhs, _p(A), ld, n * ld, _p(B), ldb, nrhs * ldb,
                                 _p(info), count, nthreads or os.cpu_count())
    return A, B, info
