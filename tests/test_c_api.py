"""The drop-in path from C: tests/c_api/dgesv_c.c is a MAGMA-style C program
(magma_init, magma_queue_create, magma_malloc / _cpu / _pinned,
magma_setmatrix, magma_dset_pointer, magma_ddisplace_pointers,
magma_iset_pointer, magma_dgesv_batched, magma_getvector, magma_getmatrix,
magma_queue_destroy, magma_finalize) compiled against include/ with gcc as C99
and linked to libmagma_amd.so.

CPU: the program compiles and links (the headers are valid C and every symbol
it calls is exported).  GPU: it runs, and its pivots, info, factors and
solutions are bit-exact against the oracle's netlib-order dgetrf + dgetrs on
the same inputs, including a singular matrix (info > 0, B left unchanged)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c_api", "dgesv_c.c")
LIBDIR = os.path.join(ROOT, "magma_amd")


def _build(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libmagma_amd.so")):
        pytest.skip("libmagma_amd.so is not built")
    exe = str(tmp_path / "dgesv_c")
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-O2", "-I" + os.path.join(ROOT, "include"), SRC,
           "-o", exe, "-L" + LIBDIR, "-lmagma_amd", "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath," + LIBDIR, "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    return exe


def test_c_program_compiles_and_links(tmp_path):
    exe = _build(tmp_path)
    assert os.access(exe, os.X_OK)


@pytest.mark.gpu
@pytest.mark.parametrize("n,nrhs,count", [(32, 1, 64), (100, 3, 12), (256, 1, 6)])
def test_c_program_dgesv_bitexact(tmp_path, oracle, n, nrhs, count):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from magma_amd import gen
    exe = _build(tmp_path)
    A = gen.general_batch(70 + n, count, n).copy()
    A[count // 2][3, :] = 0.0                                # column 4 zero (storage [col, row]): info = 4, B unchanged
    B = gen.rhs_batch(70 + n, count, n, nrhs)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as f:
        np.array([count, n, nrhs], dtype=np.int32).tofile(f)
        np.ascontiguousarray(A, dtype=np.float64).tofile(f)
        np.ascontiguousarray(B, dtype=np.float64).tofile(f)
    print(f"C program dgesv n={n} nrhs={nrhs} count={count}", flush=True)
    r = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(fout, "rb").read()
    off = 0
    ret = np.frombuffer(raw, np.int32, 1, off)[0]; off += 4
    info = np.frombuffer(raw, np.int32, count, off); off += 4 * count
    ipiv = np.frombuffer(raw, np.int32, count * n, off).reshape(count, n); off += 4 * count * n
    LU = np.frombuffer(raw, np.float64, count * n * n, off).reshape(count, n, n); off += 8 * count * n * n
    X = np.frombuffer(raw, np.float64, count * n * nrhs, off).reshape(count, nrhs, n)
    assert ret == 0
    LUo, ipo, Xo, info_o = oracle.batch_dgesv(A, B)
    assert np.array_equal(info, info_o) and info[count // 2] == 4
    assert np.array_equal(ipiv, ipo)
    assert np.array_equal(LU, LUo)
    for s in range(count):
        if info_o[s] == 0:
            assert np.array_equal(X[s], Xo[s]), s
        else:
            assert np.array_equal(X[s], B[s]), s             # dgesv leaves B as given for a singular factor
