"""GPU parity: magma_dpotrf/dpotrs/dposv_batched against the CPU oracle.

Bar: info bit-exact; the Cholesky factor bit-exact (same update order and
fused multiply-adds as the oracle); the upper triangle of A untouched; the
solves to rounding (the fused n <= 64 posv scales by 1/l_kk where dtrsm
divides, and its L' sweep associates in the other order), checked against the
oracle's solution, by the tester residual < 30 and the factor residual
||A - LL'||/(n||A||eps) < 30.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 7, 8, 15, 16, 17, 31, 32, 33, 47, 64, 65, 100, 128, 129, 256, 300, 384, 385, 512, 513, 1024]


def same(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def spd(count, n, seed):
    from magma_amd import gen
    return gen.spd_batch(seed, count, n)


def lower(A):
    # storage (cols, rows): element (i, j) at [j, i]; lower triangle = i >= j
    n = A.shape[-1]
    mask = np.tril(np.ones((n, n), bool)).T      # [j, i] with i >= j
    return np.where(mask, A, 0.0)


@pytest.mark.parametrize("n", SIZES)
def test_potrf_bitexact(queue, oracle, n):
    import magma_amd as ma
    count = 200 if n <= 64 else (30 if n <= 300 else (6 if n <= 513 else 3))
    A = spd(count, n, seed=n)
    # poison the strict upper triangle: the routine must neither read nor write it
    Ain = A.copy()
    up = ~np.tril(np.ones((n, n), bool)).T
    Ain[:, up] = 7.0e300
    print(f"potrf n={n} count={count}", flush=True)
    dA = ma.DeviceBatch(Ain)
    ret, info = ma.dpotrf_batched(dA, queue)
    assert ret == 0
    L = dA.numpy()
    assert np.all(info.cpu().numpy() == 0)
    assert np.all(L[:, up] == 7.0e300)
    for s in range(count):
        Lo, info_o = oracle.dpotrf(A[s])
        assert info_o == 0
        assert same(lower(L[s]), lower(Lo))
    for s in range(0, count, max(1, count // 4)):
        assert oracle.chol_error(A[s], lower(L[s])) < 30


@pytest.mark.parametrize("n", [5, 40, 64, 150, 385, 1024])
def test_potrf_info(queue, oracle, n):
    import magma_amd as ma
    count = 12
    A = spd(count, n, seed=3)
    for s in range(count):
        k = (s * 5 + 1) % n
        if s % 3 == 0:
            A[s, k, k] = -1.0e6                    # not positive definite at k
        elif s % 3 == 1:
            A[s, k, k] = np.nan
    dA = ma.DeviceBatch(A)
    ret, info = ma.dpotrf_batched(dA, queue)
    info = info.cpu().numpy()
    L = dA.numpy()
    for s in range(count):
        Lo, info_o = oracle.dpotrf(A[s])
        assert info[s] == info_o
        k = info_o - 1 if info_o > 0 else n
        # the leading (info-1) x (info-1) block is final
        assert same(lower(L[s])[:k, :k], lower(Lo)[:k, :k])


@pytest.mark.parametrize("n,nrhs", [(1, 1), (8, 2), (32, 4), (64, 4), (64, 1), (65, 3), (128, 8), (256, 11)])
def test_posv(queue, oracle, n, nrhs):
    import magma_amd as ma
    from magma_amd import gen
    count = 100 if n <= 64 else 20
    A = spd(count, n, seed=200 + n)
    B = gen.rhs_batch(200 + n, count, n, nrhs)
    print(f"posv n={n} nrhs={nrhs}", flush=True)
    dA = ma.DeviceBatch(A)
    dB = ma.DeviceBatch(B)
    ret, info = ma.dposv_batched(dA, dB, queue)
    assert ret == 0
    assert np.all(info.cpu().numpy() == 0)
    X = dB.numpy()
    Lo, Xo, infoo = oracle.batch_dposv(A, B)
    assert same(lower(dA.numpy()), lower(Lo))
    assert np.allclose(X, Xo, rtol=1e-9, atol=1e-14)
    for s in range(0, count, max(1, count // 5)):
        assert oracle.solve_error(A[s], X[s], B[s]) < 30


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n,nrhs", [(16, 3), (64, 4), (100, 2)])
def test_potrs_both_triangles(queue, oracle, uplo, n, nrhs):
    import magma_amd as ma
    from magma_amd import gen
    count = 16
    A = spd(count, n, seed=9)
    B = gen.rhs_batch(9, count, n, nrhs)
    F = np.stack([oracle.dpotrf(A[s], uplo=uplo.encode())[0] for s in range(count)])
    dF = ma.DeviceBatch(F)
    dB = ma.DeviceBatch(B)
    u = ma.MagmaLower if uplo == "L" else ma.MagmaUpper
    assert ma.dpotrs_batched(dF, dB, queue, uplo=u) == 0
    X = dB.numpy()
    for s in range(count):
        Xo = oracle.dpotrs(F[s], B[s], uplo=uplo.encode())
        assert np.allclose(X[s], Xo, rtol=1e-10, atol=1e-14)
        assert oracle.solve_error(A[s], X[s], B[s]) < 30


def test_cfg2_full_size(queue, oracle):
    """20000 SPD matrices n=64 (seed 1), nrhs=4: info exact, factor bit-exact."""
    import magma_amd as ma
    from magma_amd import gen
    count, n = 20000, 64
    A = gen.spd_batch(1, count, n)
    B = gen.rhs_batch(1, count, n, 4)
    dA = ma.DeviceBatch(A)
    dB = ma.DeviceBatch(B)
    ret, info = ma.dposv_batched(dA, dB, queue)
    assert ret == 0
    Lo, Xo, infoo = oracle.batch_dposv(A, B)
    assert np.array_equal(info.cpu().numpy(), infoo)
    assert same(lower(dA.numpy()), lower(Lo))
    assert np.allclose(dB.numpy(), Xo, rtol=1e-9, atol=1e-14)


def test_potrf_argument_errors(queue):
    import magma_amd as ma
    L = ma.load()
    assert L.magma_dpotrf_batched(0, 4, 0, 4, 0, 1, queue.handle) == -1             # no such uplo
    assert L.magma_dpotrf_batched(ma.MagmaLower, -1, 0, 4, 0, 1, queue.handle) == -2
    assert L.magma_dpotrf_batched(ma.MagmaLower, 4, 0, 3, 0, 1, queue.handle) == -4


# ------------------------------------------------------------------ Upper ---
def upper(A):
    # storage [j, i] = A(i, j); upper triangle = i <= j
    n = A.shape[-1]
    mask = np.triu(np.ones((n, n), bool)).T
    return np.where(mask, A, 0.0)


@pytest.mark.parametrize("n", [1, 2, 7, 32, 33, 64, 65, 100, 129, 256, 300, 600])
def test_potrf_upper_bitexact(queue, oracle, n):
    """MagmaUpper: A = U'U from the upper triangle, bit-exact against the
    oracle's upper dpotrf (the mirror image of the lower loop order); the
    strict lower triangle is neither read nor written."""
    import magma_amd as ma
    count = 40 if n <= 64 else (12 if n <= 300 else 3)
    A = spd(count, n, seed=300 + n)
    Ain = A.copy()
    lo = ~np.triu(np.ones((n, n), bool)).T                 # strict lower in [j, i] storage
    Ain[:, lo] = 7.0e300
    print(f"potrf upper n={n} count={count}", flush=True)
    dA = ma.DeviceBatch(Ain)
    ret, info = ma.dpotrf_batched(dA, queue, uplo=ma.MagmaUpper)
    assert ret == 0
    U = dA.numpy()
    assert np.all(info.cpu().numpy() == 0)
    assert np.all(U[:, lo] == 7.0e300)
    for s in range(count):
        Uo, info_o = oracle.dpotrf(A[s], uplo=b"U")
        assert info_o == 0
        assert same(upper(U[s]), upper(Uo)), s
    assert oracle.chol_error(A[0], upper(U[0]), uplo=b"U") < 30


@pytest.mark.parametrize("n", [5, 64, 150])
def test_potrf_upper_info(queue, oracle, n):
    import magma_amd as ma
    count = 6
    A = spd(count, n, seed=n + 1).copy()
    A[1][n // 2, n // 2] = -1.0                                       # not positive before any update
    A[2][min(3, n - 1), min(3, n - 1)] = np.nan
    A[3] = -A[3]
    A[4][0, n - 1] = A[4][n - 1, 0] = 1e6                             # indefinite after elimination
    dA = ma.DeviceBatch(A)
    print(f"potrf upper info n={n}", flush=True)
    ret, info = ma.dpotrf_batched(dA, queue, uplo=ma.MagmaUpper)
    assert ret == 0
    U = dA.numpy()
    for s in range(count):
        Uo, info_o = oracle.dpotrf(A[s], uplo=b"U")
        assert info.cpu().numpy()[s] == info_o, s
        # after a failed column only the leading (info-1) x (info-1) block is
        # final (LAPACK leaves the rest unspecified), as in test_potrf_info
        k = info_o - 1 if info_o > 0 else n
        assert same(upper(U[s])[:k, :k], upper(Uo)[:k, :k]), s


@pytest.mark.parametrize("n,nrhs", [(16, 3), (64, 4), (100, 2), (300, 1)])
def test_posv_upper(queue, oracle, n, nrhs):
    import magma_amd as ma
    from magma_amd import gen
    count = 10
    A = spd(count, n, seed=40 + n).copy()
    A[3][n // 3, n // 3] = -2.0                                       # info > 0: B unchanged
    B = gen.rhs_batch(40 + n, count, n, nrhs)
    dA, dB = ma.DeviceBatch(A), ma.DeviceBatch(B)
    print(f"posv upper n={n} nrhs={nrhs}", flush=True)
    ret, info = ma.dposv_batched(dA, dB, queue, uplo=ma.MagmaUpper)
    assert ret == 0
    U, X = dA.numpy(), dB.numpy()
    inf = info.cpu().numpy()
    for s in range(count):
        Uo, info_o = oracle.dpotrf(A[s], uplo=b"U")
        assert inf[s] == info_o
        k = info_o - 1 if info_o > 0 else n
        assert same(upper(U[s])[:k, :k], upper(Uo)[:k, :k])
        if info_o:
            assert np.array_equal(X[s], B[s])
            continue
        Xo = oracle.dpotrs(Uo, B[s], uplo=b"U")
        assert np.allclose(X[s], Xo, rtol=1e-10, atol=1e-14)
        assert oracle.solve_error(A[s], X[s], B[s]) < 30


@pytest.mark.parametrize("n,nrhs", [(40, 2), (200, 1), (400, 3)])
def test_dsposv_upper_same_as_lower(queue, oracle, n, nrhs):
    """dsposv with MagmaUpper on the upper triangle gives the bits of the
    MagmaLower call on the lower triangle of the same symmetric matrix (the
    kernels see the same values), leaves A unchanged, and meets n eps."""
    import magma_amd as ma
    import torch
    from magma_amd import gen
    count = 8
    A = gen.spd_batch(60 + n, count, n)
    B = gen.rhs_batch(60 + n, count, n, nrhs)
    L = ma.load()
    res = {}
    for uplo in (ma.MagmaLower, ma.MagmaUpper):
        Ain = A.copy()
        keep = np.tril(np.ones((n, n), bool)).T if uplo == ma.MagmaLower else np.triu(np.ones((n, n), bool)).T
        Ain[:, ~keep] = 5.0e300                                        # the other triangle is not read
        dA, dB, dX = ma.DeviceBatch(Ain), ma.DeviceBatch(B), ma.DeviceBatch(np.zeros_like(B))
        info = torch.zeros(count, dtype=torch.int32, device="cuda")
        it = torch.zeros(count, dtype=torch.int32, device="cuda")
        print(f"dsposv uplo={uplo} n={n}", flush=True)
        ret = L.magma_dsposv_batched(uplo, n, nrhs, dA.ptr, dA.ld, dB.ptr, dB.ld, dX.ptr, dX.ld,
                                     info.data_ptr(), it.data_ptr(), count, queue.handle)
        assert ret == 0
        assert np.array_equal(dA.numpy(), Ain)
        res[uplo] = (dX.numpy(), it.cpu().numpy())
    Xl, itl = res[ma.MagmaLower]
    Xu, itu = res[ma.MagmaUpper]
    assert np.array_equal(itl, itu) and np.all(itl >= 0)
    assert np.array_equal(Xl, Xu)
    for s in range(count):
        assert oracle.solve_error(A[s], Xu[s], B[s]) <= 1.0


def test_potrf_vbatched_upper(queue, oracle):
    import magma_amd as ma
    sizes = [1, 17, 64, 65, 130, 0, 300, 9]
    mats = [spd(1, n, seed=80 + i)[0] if n else np.zeros((0, 0)) for i, n in enumerate(sizes)]
    mats[7] = mats[7].copy()
    mats[7][4, 4] = -1.0
    V = ma.VarBatch(mats)
    info = ma.dev_ints(np.full(len(sizes), 77))
    L = ma.load()
    print(f"potrf_vbatched upper sizes={sizes}", flush=True)
    assert L.magma_dpotrf_vbatched(ma.MagmaUpper, V.d_cols.data_ptr(), V.ptr, V.d_ld.data_ptr(),
                                   info.data_ptr(), len(sizes), queue.handle) == 0
    out = V.matrices()
    inf = info.cpu().numpy()[:len(sizes)]
    for i, n in enumerate(sizes):
        if n == 0:
            continue
        Uo, info_o = oracle.dpotrf(mats[i], uplo=b"U")
        assert inf[i] == info_o, i
        assert same(upper(out[i]), upper(Uo)), i
        lo = ~np.triu(np.ones((n, n), bool)).T
        assert np.array_equal(out[i][lo], mats[i][lo]), i
