"""GPU parity above order 2048 (up to 8192): the blocked factor kernels with
narrowed panels (16 rows of the panel column per thread and two columns up to
4096, 32 rows and one column up to 8192), the solves at those orders, and the
largest vbatched size class, against the oracle.  The reference's batched
drivers print a warning above 2048 and factor anyway
(src/zgetrf_batched.cpp:114-121, src/zpotrf_batched.cpp:32-39).

Bar as at the smaller orders: ipiv and info bit-exact, factors bit-exact, the
no-transpose and forward solves bit-exact, tester residuals < 30.  Orders above
8192 are refused with MAGMA_ERR_NOT_SUPPORTED.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def same(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


@pytest.mark.parametrize("m,n,count", [(2100, 2100, 2), (4100, 24, 3), (30, 2500, 2)])
def test_getrf_getrs_above_2048(queue, oracle, m, n, count):
    import magma_amd as ma
    from magma_amd import gen
    A = gen.uniform_pm1(60 + m, np.arange(count), n * m).reshape(count, n, m)
    print(f"getrf m={m} n={n} count={count}", flush=True)
    dA = ma.DeviceBatch(A)
    ret, ipiv, info = ma.dgetrf_batched(dA, queue, m=m, n=n)
    assert ret == 0
    LU, ip, inf = dA.numpy(), ipiv.numpy(), info.cpu().numpy()
    for s in range(count):
        LUo, ipo, info_o = oracle.dgetrf(A[s], m=m, n=n)
        assert inf[s] == info_o
        assert np.array_equal(ip[s], ipo)
        assert same(LU[s], LUo)
    if m == n:
        assert oracle.lu_error(A[0], LU[0], ip[0]) < 30
        B = gen.rhs_batch(61, count, n, 3)
        print(f"getrs n={n} nrhs=3", flush=True)
        dB = ma.DeviceBatch(B)
        assert ma.dgetrs_batched(dA, ipiv, dB, queue) == 0
        X = dB.numpy()
        for s in range(count):
            assert same(X[s], oracle.dgetrs(LU[s], ip[s], B[s]))
        assert oracle.solve_error(A[0], X[0], B[0]) < 30


def _dominant_spd(seed, n):
    """Symmetric, off-diagonal entries in [-1, 1], diagonal n: SPD by
    diagonal dominance (cheap to make at large n)."""
    rng = np.random.default_rng(seed)
    S = rng.uniform(-1.0, 1.0, (n, n))
    A = np.triu(S, 1)
    A = A + A.T
    A[np.diag_indices(n)] = float(n)
    return A


def test_potrf_posv_above_2048(queue, oracle):
    """n = 2200 (two columns per panel): a full factor and solve, and a
    planted negative pivot; n = 4100 (one column per panel) stopped by a
    negative pivot at column 11, so that the factor work stays small.  The
    strict upper triangle is never read or written."""
    import magma_amd as ma
    from magma_amd import gen
    for n, stop in [(2200, None), (2200, 1500), (4100, 10)]:
        A = _dominant_spd(70 + n + (stop or 0), n)
        if stop is not None:
            A[stop, stop] = -1.0
        Ain = A[None].copy()
        up = ~np.tril(np.ones((n, n), bool)).T
        Ain[:, up] = 7.0e300
        print(f"potrf n={n} negative pivot at {stop}", flush=True)
        dA = ma.DeviceBatch(Ain)
        ret, info = ma.dpotrf_batched(dA, queue)
        assert ret == 0
        L = dA.numpy()[0]
        inf = int(info.cpu().numpy()[0])
        assert np.all(L[up] == 7.0e300)
        Lo, info_o = oracle.dpotrf(A)
        assert inf == info_o
        assert inf == (0 if stop is None else stop + 1)
        low = ~up
        k = n if stop is None else stop                 # the leading (info-1) x (info-1) block is final
        assert same(np.where(low, L, 0.0)[:k, :k], np.where(low, Lo, 0.0)[:k, :k])
        if stop is None:
            assert oracle.chol_error(A, np.where(low, L, 0.0)) < 30
            B = gen.rhs_batch(71, 1, n, 2)
            print(f"potrs n={n} nrhs=2", flush=True)
            dB = ma.DeviceBatch(B)
            assert ma.dpotrs_batched(dA, dB, queue) == 0
            X = dB.numpy()[0]
            assert oracle.solve_error(A, X, B[0]) < 30


def test_getrf_vbatched_largest_class(queue, oracle):
    """Orders 2049..8192 form the last size class; checked and unchecked
    entry points give the fixed-size bits."""
    import magma_amd as ma
    from magma_amd import gen
    sizes = [7, 2100, 300, 0, 65]
    mats = [gen.uniform_pm1(80, [i], n * n).reshape(n, n) for i, n in enumerate(sizes)]
    L = ma.load()
    for nocheck in (False, True):
        V = ma.VarBatch(mats)
        ipiv = ma.VarInts([max(n, 1) for n in sizes])
        info = ma.dev_ints(np.zeros(len(mats)))
        print(f"getrf_vbatched sizes={sizes} nocheck={nocheck}", flush=True)
        if nocheck:
            ret = L.magma_dgetrf_vbatched_max_nocheck(V.d_rows.data_ptr(), V.d_cols.data_ptr(), max(sizes),
                                                      V.ptr, V.d_ld.data_ptr(), ipiv.ptr, info.data_ptr(),
                                                      V.count, queue.handle)
            queue.sync()
        else:
            ret = L.magma_dgetrf_vbatched(V.d_rows.data_ptr(), V.d_cols.data_ptr(), V.ptr, V.d_ld.data_ptr(),
                                          ipiv.ptr, info.data_ptr(), V.count, queue.handle)
        assert ret == 0
        got = V.matrices()
        info_h = info.cpu().numpy()
        for i, n in enumerate(sizes):
            if n == 0:
                assert info_h[i] == 0
                continue
            LUo, ipo, info_o = oracle.dgetrf(mats[i])
            assert info_h[i] == info_o, i
            assert np.array_equal(ipiv.numpy(i, n), ipo), i
            assert same(got[i], LUo), i


def test_dsgesv_vbatched_largest_class(queue, oracle):
    """An order above 2048 in a mixed-precision batch takes the FP64 path
    (iter = -1) and gets the plain factor and solve bits."""
    import magma_amd as ma
    from magma_amd import gen
    sizes = [2100, 40]
    mats = [gen.uniform_pm1(90, [i], n * n).reshape(n, n) for i, n in enumerate(sizes)]
    rhs = [gen.rhs_batch(91 + i, 1, n, 1)[0] for i, n in enumerate(sizes)]
    V = ma.VarBatch(mats)
    R = ma.VarBatch(rhs)
    Xv = ma.VarBatch([np.zeros_like(b) for b in rhs])
    ipiv = ma.VarInts(sizes)
    info = ma.dev_ints(np.full(2, 77))
    it = ma.dev_ints(np.full(2, 99))
    print(f"dsgesv_vbatched sizes={sizes}", flush=True)
    ret = ma.load().magma_dsgesv_vbatched(V.d_cols.data_ptr(), R.d_cols.data_ptr(), V.ptr, V.d_ld.data_ptr(),
                                          ipiv.ptr, R.ptr, R.d_ld.data_ptr(), Xv.ptr, Xv.d_ld.data_ptr(),
                                          info.data_ptr(), it.data_ptr(), 2, queue.handle)
    assert ret == 0
    X = Xv.matrices()
    A = V.matrices()
    itv = it.cpu().numpy()[:2]
    infv = info.cpu().numpy()[:2]
    assert itv[0] == -1 and infv[0] == 0
    LUo, ipo, _ = oracle.dgetrf(mats[0])
    assert np.array_equal(ipiv.numpy(0, 2100), ipo)
    assert same(A[0], LUo)
    assert same(X[0], oracle.dgetrs(LUo, ipo, rhs[0]))
    assert 0 <= itv[1] <= 10 and infv[1] == 0
    assert oracle.solve_error(mats[1], X[1], rhs[1]) <= 1.0


def test_above_8192_not_supported(queue):
    """Orders above 8192 are refused before any work (no allocation of the
    matrices needed: the checks read only the sizes)."""
    import magma_amd as ma
    from magma_amd import lib as ma_lib
    L = ma.load()
    q = queue.handle
    big = ma.dev_ints([4, 9000, 4])
    info = ma.dev_ints(np.zeros(3))
    assert L.magma_dgetrf_vbatched_max_nocheck(big.data_ptr(), big.data_ptr(), 9000, 0, big.data_ptr(), 0,
                                               info.data_ptr(), 3, q) == ma_lib.MAGMA_ERR_NOT_SUPPORTED
    assert L.magma_dgetrf_vbatched(big.data_ptr(), big.data_ptr(), 0, big.data_ptr(), 0, info.data_ptr(), 3,
                                   q) == ma_lib.MAGMA_ERR_NOT_SUPPORTED
