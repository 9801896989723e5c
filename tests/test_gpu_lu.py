"""GPU parity: magma_dgetrf/dgetrs/dgesv_batched (libmagma_amd.so, C-ABI)
against the CPU oracle on the same seeded inputs.

Bar: ipiv and info bit-exact; the LU factors and the no-transpose solutions
bit-exact as well (the kernels apply each element's updates in the oracle's
order with the same fused multiply-adds); tester residuals < 30.  The
transposed solve associates its sums differently from dtrsm and is checked
to rounding (residual < 30, relative difference < 1e-10).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 5, 8, 9, 13, 16, 17, 24, 31, 32, 33, 40, 48, 63, 64, 65, 96, 100, 128, 129, 200, 256, 257, 300,
         384, 512, 513, 600]


def _batch(count, n, seed, m=None, ld=None):
    from magma_amd import gen
    m = n if m is None else m
    A = gen.uniform_pm1(seed, np.arange(count), n * m).reshape(count, n, m)
    return A


def same(a, b):
    """bitwise-equal up to the sign of zero, NaN == NaN"""
    a = np.asarray(a); b = np.asarray(b)
    return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def run_getrf(queue, A, m, n, ld=None):
    import magma_amd as ma
    print(f"getrf m={m} n={n} count={A.shape[0]} ld={ld}", flush=True)
    dA = ma.DeviceBatch(A, ld=ld)
    ret, ipiv, info = ma.dgetrf_batched(dA, queue, m=m, n=n)
    assert ret == 0
    return dA.numpy(), ipiv.numpy(), info.cpu().numpy()


@pytest.mark.parametrize("n", SIZES)
def test_getrf_bitexact(queue, oracle, n):
    count = 300 if n <= 64 else 40
    A = _batch(count, n, seed=n)
    LU, ipiv, info = run_getrf(queue, A, n, n)
    LUo, ipivo, infoo = oracle.batch_dgetrf(A)
    assert np.array_equal(info, infoo)
    assert np.array_equal(ipiv, ipivo)
    assert same(LU, LUo)
    for s in range(0, count, max(1, count // 5)):
        assert oracle.lu_error(A[s], LU[s], ipiv[s]) < 30


@pytest.mark.parametrize("m,n", [(20, 12), (12, 20), (64, 40), (40, 64), (150, 90), (90, 150), (10, 250),
                                 (250, 10), (400, 260), (260, 400), (512, 33), (33, 512)])
def test_getrf_rectangular(queue, oracle, m, n):
    count = 50
    A = _batch(count, n, seed=7, m=m)
    LU, ipiv, info = run_getrf(queue, A, m, n)
    for s in range(count):
        LUo, ipo, info_o = oracle.dgetrf(A[s], m=m, n=n)
        assert info[s] == info_o
        assert np.array_equal(ipiv[s], ipo)
        assert same(LU[s], LUo)


@pytest.mark.parametrize("n,ld", [(32, 40), (64, 70), (100, 128), (200, 256), (300, 301)])
def test_getrf_padded_ld(queue, oracle, n, ld):
    A = _batch(20, n, seed=3)
    LU, ipiv, info = run_getrf(queue, A, n, n, ld=ld)
    LUo, ipivo, infoo = oracle.batch_dgetrf(A)
    assert np.array_equal(ipiv, ipivo) and same(LU, LUo)


@pytest.mark.parametrize("n", [6, 32, 50, 130, 300])
def test_getrf_singular_ties_nan(queue, oracle, n):
    rng = np.random.default_rng(n)
    mats = []
    # integer entries (exact ties), zero columns, duplicated rows, a NaN
    for t in range(24):
        A = rng.integers(-2, 3, size=(n, n)).astype(float)
        if t % 4 == 1:
            A[:, t % n] = 0.0                      # zero column -> info
        if t % 4 == 2:
            A[(t + 1) % n, :] = A[t % n, :]        # singular, exact ties
        if t % 4 == 3:
            A[(t * 7) % n, (t * 3) % n] = np.nan
        mats.append(A.T.copy())
    A = np.stack(mats)
    LU, ipiv, info = run_getrf(queue, A, n, n)
    LUo, ipivo, infoo = oracle.batch_dgetrf(A)
    assert np.array_equal(info, infoo)
    assert np.array_equal(ipiv, ipivo)
    assert same(LU, LUo)


def test_getrf_zero_and_identity(queue, oracle):
    n = 40
    A = np.stack([np.eye(n), np.zeros((n, n)), np.eye(n)[::-1].T.copy()])
    LU, ipiv, info = run_getrf(queue, A, n, n)
    assert list(info) == [0, 1, 0]
    assert list(ipiv[0]) == list(range(1, n + 1))
    assert list(ipiv[1]) == list(range(1, n + 1))
    LUo, ipivo, infoo = oracle.batch_dgetrf(A)
    assert np.array_equal(ipiv, ipivo) and same(LU, LUo)


@pytest.mark.parametrize("n,nrhs", [(1, 1), (5, 3), (16, 1), (32, 1), (32, 5), (48, 2), (64, 4), (65, 1),
                                    (128, 3), (256, 1)])
def test_gesv_bitexact(queue, oracle, n, nrhs):
    import magma_amd as ma
    from magma_amd import gen
    count = 200 if n <= 64 else 30
    A = _batch(count, n, seed=1000 + n)
    B = gen.rhs_batch(1000 + n, count, n, nrhs)
    print(f"gesv n={n} nrhs={nrhs}", flush=True)
    dA = ma.DeviceBatch(A)
    dB = ma.DeviceBatch(B)
    ret, ipiv, info = ma.dgesv_batched(dA, dB, queue)
    assert ret == 0
    LUo, ipivo, Xo, infoo = oracle.batch_dgesv(A, B)
    assert np.array_equal(info.cpu().numpy(), infoo)
    assert np.array_equal(ipiv.numpy(), ipivo)
    assert same(dA.numpy(), LUo)
    X = dB.numpy()
    assert same(X, Xo)
    for s in range(0, count, max(1, count // 5)):
        assert oracle.solve_error(A[s], X[s], B[s]) < 30


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("n,nrhs", [(7, 2), (32, 1), (64, 3), (100, 9), (256, 2)])
def test_getrs(queue, oracle, n, nrhs, trans):
    import magma_amd as ma
    from magma_amd import gen
    count = 30
    A = _batch(count, n, seed=50 + n)
    B = gen.rhs_batch(50 + n, count, n, nrhs)
    print(f"getrs n={n} nrhs={nrhs} trans={trans}", flush=True)
    dA = ma.DeviceBatch(A)
    ret, ipiv, info = ma.dgetrf_batched(dA, queue)
    dB = ma.DeviceBatch(B)
    t = ma.MagmaNoTrans if trans == "N" else ma.MagmaTrans
    assert ma.dgetrs_batched(dA, ipiv, dB, queue, trans=t) == 0
    X = dB.numpy()
    LU = dA.numpy()
    for s in range(count):
        Xo = oracle.dgetrs(LU[s], ipiv.numpy()[s], B[s], trans=trans.encode())
        if trans == "N":
            assert same(X[s], Xo)
        else:
            assert np.allclose(X[s], Xo, rtol=1e-10, atol=1e-12)
        At = A[s] if trans == "N" else np.ascontiguousarray(A[s].T)
        assert oracle.solve_error(At, X[s], B[s]) < 30


def test_cfg1_full_size(queue, oracle):
    """10000 matrices n=32 (seed 0), one RHS: every pivot, info and solution
    entry equal to the oracle's."""
    import magma_amd as ma
    from magma_amd import gen
    count, n = 10000, 32
    A = gen.general_batch(0, count, n)
    B = gen.rhs_batch(0, count, n, 1)
    dA = ma.DeviceBatch(A)
    dB = ma.DeviceBatch(B)
    ret, ipiv, info = ma.dgesv_batched(dA, dB, queue)
    assert ret == 0
    LUo, ipivo, Xo, infoo = oracle.batch_dgesv(A, B)
    assert np.array_equal(info.cpu().numpy(), infoo)
    assert np.array_equal(ipiv.numpy(), ipivo)
    assert same(dB.numpy(), Xo)


def test_argument_errors(queue):
    import magma_amd as ma
    L = ma.load()
    q = queue.handle
    assert L.magma_dgetrf_batched(-1, 4, 0, 4, 0, 0, 1, q) == -1
    assert L.magma_dgetrf_batched(4, -1, 0, 4, 0, 0, 1, q) == -2
    assert L.magma_dgetrf_batched(4, 4, 0, 3, 0, 0, 1, q) == -4
    assert L.magma_dgetrs_batched(0, 4, 1, 0, 4, 0, 0, 4, 1, q) == -1
    assert L.magma_dgetrs_batched(ma.MagmaNoTrans, 4, -1, 0, 4, 0, 0, 4, 1, q) == -3
    assert L.magma_dgetrs_batched(ma.MagmaNoTrans, 4, 1, 0, 4, 0, 0, 3, 1, q) == -8
    assert L.magma_dgesv_batched(4, 1, 0, 2, 0, 0, 4, 0, 1, q) == -4
    assert L.magma_dgesv_batched(4, 1, 0, 4, 0, 0, 2, 0, 1, q) == -6
    # quick returns touch nothing
    assert L.magma_dgetrf_batched(0, 0, 0, 1, 0, 0, 5, q) == 0


@pytest.mark.parametrize("n", [5, 100, 300])
def test_getrs_zero_skip_like_dtrsm(queue, oracle, n):
    """Reference dtrsm skips the whole column update for a zero solution entry
    (NoTrans), so an Inf below that pivot in L must not turn the solution into
    NaN: with L(1,0) = 0 and b(1) = 0, x(1) is zero at its pivot and L(3,1) = Inf
    is never used.  Bitwise equal to the oracle's dgetrs; covers the one-row-
    per-thread sweep (n <= 256) and the general one."""
    import magma_amd as ma
    rng = np.random.default_rng(n)
    count = 3
    LU = rng.uniform(-1, 1, (count, n, n)) / n
    LU[:, np.arange(n), np.arange(n)] = rng.uniform(1, 2, (count, n))
    LU[:, 0, 1] = 0.0                      # L(1,0) = 0  (storage [j, i] = A(i, j))
    LU[:, 1, 3] = np.inf                   # L(3,1) = Inf
    B = rng.uniform(-1, 1, (count, 2, n))
    B[:, :, 1] = 0.0
    dA = ma.DeviceBatch(LU)
    ipiv = ma.IntBatch(count, n)
    ipiv.t[:] = ma._torch().arange(1, n + 1, dtype=ma._torch().int32, device="cuda")
    dB = ma.DeviceBatch(B)
    print(f"getrs zero skip n={n}", flush=True)
    assert ma.dgetrs_batched(dA, ipiv, dB, queue) == 0
    X = dB.numpy()
    for s in range(count):
        Xo = oracle.dgetrs(LU[s], np.arange(1, n + 1, dtype=np.int32), B[s])
        assert np.all(np.isfinite(Xo))
        assert same(X[s], Xo), s
