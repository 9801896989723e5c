"""magma_amd -- MI355X-native batched small-matrix factor-and-solve.

The product is libmagma_amd.so (C-ABI, MAGMA's magma_*_batched API, HIP
kernels for gfx950).  This package holds its sources (csrc/), a ctypes binding
(lib.py), deterministic input generators (gen.py) and a thin torch-based helper
layer (this module) that owns device memory and pointer arrays for Python
callers.  torch is plumbing here: allocation, copies and the stream; every
factorization and solve runs in the library's kernels.
"""
import ctypes

import numpy as np

from . import lib as _libmod
from .lib import (MagmaConjTrans, MagmaLeft, MagmaLower, MagmaNoTrans, MagmaNonUnit,  # noqa: F401
                  MagmaRight, MagmaTrans, MagmaUnit, MagmaUpper)

__all__ = ["load", "Queue", "DeviceBatch", "dgetrf_batched", "dgetrs_batched",
           "dgesv_batched", "dpotrf_batched", "dpotrs_batched", "dposv_batched"]


def load():
    return _libmod.load()


def _torch():
    import torch
    return torch


class Queue:
    """magma_queue_t on torch's current HIP stream (so torch copies and the
    library's kernels are ordered on one stream)."""

    def __init__(self, device=0, stream=None):
        torch = _torch()
        self.lib = load()
        self.device = device
        if stream is None:
            stream = torch.cuda.current_stream(device)
        self.stream = stream
        q = ctypes.c_void_p()
        self.lib.magma_queue_create_from_hip_internal(
            device, ctypes.c_void_p(stream.cuda_stream), ctypes.byref(q), b"Queue", b"__init__.py", 0)
        if not q.value:
            raise RuntimeError("magma_queue_create_from_hip failed")
        self.handle = q

    def sync(self):
        self.lib.magma_queue_sync_internal(self.handle, b"sync", b"__init__.py", 0)

    def close(self):
        if self.handle is not None and self.handle.value:
            self.lib.magma_queue_destroy_internal(self.handle, b"close", b"__init__.py", 0)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ptr_array(t, count, stride_elems, offset_elems=0):
    """Device array of `count` pointers t.data_ptr() + (offset + i*stride) elements."""
    torch = _torch()
    base = t.data_ptr() + offset_elems * t.element_size()
    ptrs = base + torch.arange(count, dtype=torch.int64) * (stride_elems * t.element_size())
    return ptrs.to(t.device)


class DeviceBatch:
    """`count` column-major matrices rows x cols with leading dimension ld,
    stored back to back (stride ld*cols) in one torch tensor, plus the device
    pointer array the C API takes."""

    def __init__(self, host, ld=None, device="cuda", dtype=None):
        torch = _torch()
        host = np.asarray(host)
        # host: (count, cols, rows) column-major storage
        count, cols, rows = host.shape
        ld = rows if ld is None else ld
        if ld < rows:
            raise ValueError("ld < rows")
        dt = dtype or torch.float64
        buf = torch.zeros((count, cols, ld), dtype=dt)
        buf[:, :, :rows] = torch.from_numpy(np.ascontiguousarray(host)).to(dt)
        self.t = buf.to(device)
        self.count, self.rows, self.cols, self.ld = count, rows, cols, ld
        self.ptrs = ptr_array(self.t, count, ld * cols)

    def numpy(self):
        return self.t[:, :, : self.rows].cpu().numpy()

    @property
    def ptr(self):
        return self.ptrs.data_ptr()


class IntBatch:
    """`count` int32 vectors of length `length`, back to back, with a pointer array."""

    def __init__(self, count, length, device="cuda"):
        torch = _torch()
        self.t = torch.zeros((count, max(length, 1)), dtype=torch.int32, device=device)
        self.ptrs = ptr_array(self.t, count, max(length, 1))
        self.length = length

    def numpy(self):
        return self.t[:, : self.length].cpu().numpy()

    @property
    def ptr(self):
        return self.ptrs.data_ptr()


class VarBatch:
    """Matrices of different sizes (rows_i x cols_i, leading dimension ld_i =
    max(rows_i, 1)) packed back to back in one device buffer, with the device
    pointer array and device size arrays (length count + 1, the vbatched
    convention) the C API takes."""

    def __init__(self, mats, device="cuda"):
        torch = _torch()
        self.count = len(mats)
        self.rows = np.array([m.shape[1] if m.ndim == 2 else 0 for m in mats], dtype=np.int32)
        self.cols = np.array([m.shape[0] if m.ndim == 2 else 0 for m in mats], dtype=np.int32)
        self.ld = np.maximum(self.rows, 1).astype(np.int32)
        sizes = (self.ld.astype(np.int64) * np.maximum(self.cols, 1))
        self.offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        flat = np.zeros(int(sizes.sum()) if self.count else 1, dtype=np.float64)
        for i, m in enumerate(mats):
            r, c = self.rows[i], self.cols[i]
            if r and c:
                blk = np.zeros((c, self.ld[i]))
                blk[:, :r] = m
                flat[self.offsets[i]: self.offsets[i] + blk.size] = blk.ravel()
        self.t = torch.from_numpy(flat).to(device)
        base = self.t.data_ptr()
        self.ptrs = torch.from_numpy(base + self.offsets * 8).to(device)
        self.d_rows = dev_ints(self.rows, device)
        self.d_cols = dev_ints(self.cols, device)
        self.d_ld = dev_ints(self.ld, device)

    def matrices(self):
        flat = self.t.cpu().numpy()
        out = []
        for i in range(self.count):
            r, c, ld, o = self.rows[i], self.cols[i], self.ld[i], self.offsets[i]
            out.append(flat[o: o + ld * c].reshape(c, ld)[:, :r].copy() if r and c else np.zeros((c, r)))
        return out

    @property
    def ptr(self):
        return self.ptrs.data_ptr()


def dev_ints(values, device="cuda"):
    """int32 device array of len(values) + 1 (the extra slot is library scratch)."""
    torch = _torch()
    a = np.zeros(len(values) + 1, dtype=np.int32)
    a[: len(values)] = values
    return torch.from_numpy(a).to(device)


class VarInts:
    """Per-matrix int32 vectors of lengths `lengths`, with a pointer array."""

    def __init__(self, lengths, device="cuda"):
        torch = _torch()
        lengths = np.maximum(np.asarray(lengths, dtype=np.int64), 1)
        self.offsets = np.concatenate([[0], np.cumsum(lengths)[:-1]])
        self.lengths = lengths
        self.t = torch.zeros(int(lengths.sum()), dtype=torch.int32, device=device)
        self.ptrs = torch.from_numpy(self.t.data_ptr() + self.offsets * 4).to(device)

    def numpy(self, i, n):
        flat = self.t.cpu().numpy()
        return flat[self.offsets[i]: self.offsets[i] + n]

    @property
    def ptr(self):
        return self.ptrs.data_ptr()


def _vec(count, device="cuda"):
    return _torch().zeros(count, dtype=_torch().int32, device=device)


def dgetrf_batched(A, queue, m=None, n=None):
    """LU of DeviceBatch A in place.  Returns (ret, ipiv IntBatch, info tensor)."""
    m = A.rows if m is None else m
    n = A.cols if n is None else n
    ipiv = IntBatch(A.count, min(m, n))
    info = _vec(A.count)
    ret = load().magma_dgetrf_batched(m, n, A.ptr, A.ld, ipiv.ptr, info.data_ptr(), A.count, queue.handle)
    return ret, ipiv, info


def dgetrs_batched(A, ipiv, B, queue, trans=MagmaNoTrans):
    return load().magma_dgetrs_batched(trans, A.cols, B.cols, A.ptr, A.ld, ipiv.ptr, B.ptr, B.ld,
                                       A.count, queue.handle)


def dgesv_batched(A, B, queue):
    ipiv = IntBatch(A.count, A.cols)
    info = _vec(A.count)
    ret = load().magma_dgesv_batched(A.cols, B.cols, A.ptr, A.ld, ipiv.ptr, B.ptr, B.ld,
                                     info.data_ptr(), A.count, queue.handle)
    return ret, ipiv, info


def dpotrf_batched(A, queue, uplo=MagmaLower):
    info = _vec(A.count)
    ret = load().magma_dpotrf_batched(uplo, A.cols, A.ptr, A.ld, info.data_ptr(), A.count, queue.handle)
    return ret, info


def dpotrs_batched(A, B, queue, uplo=MagmaLower):
    return load().magma_dpotrs_batched(uplo, A.cols, B.cols, A.ptr, A.ld, B.ptr, B.ld,
                                       A.count, queue.handle)


def dposv_batched(A, B, queue, uplo=MagmaLower):
    info = _vec(A.count)
    ret = load().magma_dposv_batched(uplo, A.cols, B.cols, A.ptr, A.ld, B.ptr, B.ld,
                                     info.data_ptr(), A.count, queue.handle)
    return ret, info
