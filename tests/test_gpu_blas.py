"""GPU parity: magmablas_dtrsm_batched, magmablas_dtrsv_batched and
magmablas_dgemv_batched against the oracle's restatements of the reference
BLAS (dtrsm.f, dtrsv.f, dgemv.f).

Bar: bitwise equal to the oracle for every case, side, triangle, transpose,
diagonal, alpha / beta (including 0 and 1) and increment; the triangle that
must not be read (and the diagonal for diag = Unit) holds NaN, and exact zeros
in the right-hand sides exercise dtrsm's / dtrsv's zero skips.
"""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def same(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def tri_batch(rng, count, k, uplo, diag):
    """(count, k, k) storage of triangular matrices with a dominant diagonal
    and NaN where the routine must not read."""
    out = np.empty((count, k, k))
    for s in range(count):
        M = rng.uniform(-1, 1, (k, k)) / max(k, 1) ** 0.5
        np.fill_diagonal(M, rng.uniform(1, 2, k) * rng.choice([-1, 1], k))
        keep = np.triu(np.ones((k, k), bool)) if uplo == "U" else np.tril(np.ones((k, k), bool))
        if diag == "U":
            np.fill_diagonal(keep, False)
        M = np.where(keep, M, np.nan)
        out[s] = M.T                                   # storage[j, i] = A(i, j)
    return out


def rhs(rng, count, cols, rows, zeros=True):
    B = rng.uniform(-1, 1, (count, cols, rows))
    if zeros:
        B[rng.uniform(size=B.shape) < 0.1] = 0.0
    return B


def enums(side, uplo, trans, diag):
    import magma_amd as ma
    return ({"L": ma.MagmaLeft, "R": ma.MagmaRight}[side] if side else None,
            ma.MagmaUpper if uplo == "U" else ma.MagmaLower,
            ma.MagmaNoTrans if trans == "N" else ma.MagmaTrans,
            ma.MagmaNonUnit if diag == "N" else ma.MagmaUnit)


TRSM_CASES = list(itertools.product("LR", "UL", "NT", "NU"))


@pytest.mark.parametrize("side,uplo,trans,diag", TRSM_CASES)
def test_trsm_batched_bitexact(queue, oracle, side, uplo, trans, diag):
    import magma_amd as ma
    L = ma.load()
    rng = np.random.default_rng(TRSM_CASES.index((side, uplo, trans, diag)))
    es, eu, et, ed = enums(side, uplo, trans, diag)
    for (m, n), alpha in itertools.product([(1, 1), (7, 5), (33, 17), (64, 70), (100, 3), (40, 1), (64, 64)],
                                           [1.0, 0.7, 0.0]):
        count = 5
        k = m if side == "L" else n
        A = tri_batch(rng, count, k, uplo, diag)
        B = rhs(rng, count, n, m)
        dA = ma.DeviceBatch(A, ld=k + 3)
        dB = ma.DeviceBatch(B, ld=m + 1)
        print(f"trsm {side}{uplo}{trans}{diag} m={m} n={n} alpha={alpha}", flush=True)
        L.magmablas_dtrsm_batched(es, eu, et, ed, m, n, alpha, dA.ptr, dA.ld, dB.ptr, dB.ld, count, queue.handle)
        got = dB.numpy()
        for s in range(count):
            want = oracle.dtrsm(side.encode(), uplo.encode(), trans.encode(), diag.encode(), m, n, alpha, A[s], B[s])
            assert same(got[s], want), (m, n, alpha, s)


@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product("UL", "NT", "NU")))
def test_trsv_batched_bitexact(queue, oracle, uplo, trans, diag):
    import magma_amd as ma
    L = ma.load()
    rng = np.random.default_rng(40)
    _, eu, et, ed = enums(None, uplo, trans, diag)
    for n, incb in itertools.product([1, 5, 64, 100, 300, 1100], [1, 2]):
        count = 4
        A = tri_batch(rng, count, n, uplo, diag)
        X = np.full((count, 1, n * incb), 9.0)
        X[:, 0, ::incb] = rhs(rng, count, 1, n)[:, 0, :]
        dA = ma.DeviceBatch(A, ld=n + 2)
        dX = ma.DeviceBatch(X)
        print(f"trsv {uplo}{trans}{diag} n={n} incb={incb}", flush=True)
        L.magmablas_dtrsv_batched(eu, et, ed, n, dA.ptr, dA.ld, dX.ptr, incb, count, queue.handle)
        got = dX.numpy()
        for s in range(count):
            want = oracle.dtrsv(uplo.encode(), trans.encode(), diag.encode(), n, A[s], X[s, 0], incb)
            assert same(got[s, 0], want), (n, incb, s)


@pytest.mark.parametrize("trans", "NT")
def test_gemv_batched_bitexact(queue, oracle, trans):
    import magma_amd as ma
    L = ma.load()
    rng = np.random.default_rng(41)
    et = ma.MagmaNoTrans if trans == "N" else ma.MagmaTrans
    sizes = [(1, 1), (33, 17), (17, 33), (300, 70), (64, 64), (5, 130)]
    for (m, n), (alpha, beta), (incx, incy) in itertools.product(
            sizes, [(1.3, -0.4), (1.0, 0.0), (0.0, 2.0), (-0.5, 1.0)], [(1, 1), (2, 3)]):
        count = 4
        lx, ly = (n, m) if trans == "N" else (m, n)
        A = rng.uniform(-1, 1, (count, n, m))
        X = np.full((count, 1, lx * incx), 7.0)
        X[:, 0, ::incx] = rng.uniform(-1, 1, (count, lx))
        Y = np.full((count, 1, ly * incy), 7.0)
        Y[:, 0, ::incy] = rng.uniform(-1, 1, (count, ly))
        if beta == 0.0:
            Y[:, 0, ::incy] = np.nan                     # beta = 0: y is not read
        dA = ma.DeviceBatch(A, ld=m + 1)
        dX = ma.DeviceBatch(X)
        dY = ma.DeviceBatch(Y)
        print(f"gemv {trans} m={m} n={n} alpha={alpha} beta={beta} inc=({incx},{incy})", flush=True)
        L.magmablas_dgemv_batched(et, m, n, alpha, dA.ptr, dA.ld, dX.ptr, incx, beta, dY.ptr, incy, count,
                                  queue.handle)
        got = dY.numpy()
        for s in range(count):
            want = oracle.dgemv(trans.encode(), m, n, alpha, A[s], X[s, 0], beta, Y[s, 0], incx, incy)
            assert same(got[s, 0], want), (m, n, alpha, beta, incx, incy, s)


def test_blas_batched_argument_errors_leave_data(queue):
    """Invalid arguments are reported (magma_xerbla) and nothing is written."""
    import magma_amd as ma
    L = ma.load()
    A = np.eye(4)[None].repeat(2, 0)
    B = np.ones((2, 3, 4))
    dA = ma.DeviceBatch(A)
    dB = ma.DeviceBatch(B)
    q = queue.handle
    L.magmablas_dtrsm_batched(0, ma.MagmaUpper, ma.MagmaNoTrans, ma.MagmaNonUnit, 4, 3, 2.0, dA.ptr, 4, dB.ptr, 4, 2, q)
    L.magmablas_dtrsm_batched(ma.MagmaLeft, ma.MagmaUpper, ma.MagmaNoTrans, ma.MagmaNonUnit, 4, 3, 2.0, dA.ptr, 3,
                              dB.ptr, 4, 2, q)
    L.magmablas_dtrsv_batched(ma.MagmaUpper, ma.MagmaNoTrans, ma.MagmaNonUnit, 4, dA.ptr, 4, dB.ptr, 0, 2, q)
    L.magmablas_dgemv_batched(ma.MagmaNoTrans, 4, 4, 2.0, dA.ptr, 4, dB.ptr, 0, 0.0, dB.ptr, 1, 2, q)
    queue.sync()
    assert np.array_equal(dB.numpy(), B)


@pytest.mark.parametrize("n", [1, 5, 32, 64, 100, 256, 300])
def test_getri_outofplace_batched(queue, oracle, n):
    """Inverse from the getrf factors: bitwise the oracle's dgetrs applied to
    the identity (column j solves L U x = e_p(j), as the reference's column
    interchange after its two triangular solves produces), with the tester's
    residual ||I - A inv(A)|| / (n ||A|| ||inv(A)|| eps) small."""
    import magma_amd as ma
    import torch
    L = ma.load()
    rng = np.random.default_rng(n)
    count = 6
    A = rng.uniform(-1, 1, (count, n, n))
    if n > 1:
        A[count - 1][2 % n, :] = 0.0                   # a zero column: U(3,3) = 0 -> info 3
    dA = ma.DeviceBatch(A, ld=n + 1)
    ret, ipiv, info = ma.dgetrf_batched(dA, queue)
    assert ret == 0
    dI = ma.DeviceBatch(np.full((count, n, n), 5.0), ld=n + 2)
    info2 = torch.full((count,), -9, dtype=torch.int32, device="cuda")
    print(f"getri n={n}", flush=True)
    assert L.magma_dgetri_outofplace_batched(n, dA.ptr, dA.ld, ipiv.ptr, dI.ptr, dI.ld, info2.data_ptr(), count,
                                             queue.handle) == 0
    LU = dA.numpy(); piv = ipiv.numpy(); inv = dI.numpy(); inf = info2.cpu().numpy(); inf_f = info.cpu().numpy()
    for s in range(count):
        want = oracle.dgetrs(LU[s], piv[s], np.eye(n))
        assert same(inv[s], want), s
        assert inf[s] == inf_f[s], (s, inf[s], inf_f[s])
        if inf[s] == 0:
            Am = A[s].T
            R = np.eye(n) - Am @ inv[s].T
            err = np.abs(R).max() / (n * np.abs(Am).max() * np.abs(inv[s]).max() * 2.0 ** -53)
            assert err < 30, (s, err)
    if n > 1:
        assert inf[count - 1] == 3


def test_getri_argument_errors(queue):
    import magma_amd as ma
    L = ma.load()
    q = queue.handle
    assert L.magma_dgetri_outofplace_batched(-1, 0, 1, 0, 0, 1, 0, 1, q) == -1
    assert L.magma_dgetri_outofplace_batched(4, 0, 3, 0, 0, 4, 0, 1, q) == -3
    assert L.magma_dgetri_outofplace_batched(4, 0, 4, 0, 0, 3, 0, 1, q) == -6


# ------------------------------------------------------------ no pivoting ----
def _dominant(rng, count, m, n):
    A = rng.uniform(-1, 1, (count, n, m))
    k = min(m, n)
    A[:, np.arange(k), np.arange(k)] += np.sign(A[:, np.arange(k), np.arange(k)]) * (m + 1)
    return A


@pytest.mark.parametrize("m,n", [(1, 1), (5, 5), (32, 32), (64, 64), (90, 90), (100, 100), (256, 256),
                                 (70, 40), (40, 70)])
def test_getrf_nopiv_batched_bitexact(queue, oracle, m, n):
    """Bitwise the oracle's no-pivot dgetf2 (LDS-resident up to 64 KB per
    matrix, in place above), on dominant and on general matrices (the latter
    with growth, NaN and Inf compared as values)."""
    import magma_amd as ma
    import torch
    L = ma.load()
    rng = np.random.default_rng(m * 1000 + n)
    count = 6
    A = np.concatenate([_dominant(rng, count // 2, m, n), rng.uniform(-1, 1, (count - count // 2, n, m))])
    if min(m, n) > 2:
        A[count - 1][1, :] = 0.0
        A[count - 1][:, 1] = 0.0                      # second pivot exactly zero -> info 2
    dA = ma.DeviceBatch(A, ld=m + 3)
    info = torch.full((count,), -5, dtype=torch.int32, device="cuda")
    print(f"getrf_nopiv m={m} n={n}", flush=True)
    assert L.magma_dgetrf_nopiv_batched(m, n, dA.ptr, dA.ld, info.data_ptr(), count, queue.handle) == 0
    got = dA.numpy(); inf = info.cpu().numpy()
    for s in range(count):
        LUo, info_o = oracle.dgetrf_nopiv(A[s], m, n)
        assert inf[s] == info_o, (s, inf[s], info_o)
        assert same(got[s], LUo), s
    if min(m, n) > 2:
        assert inf[count - 1] == 2


@pytest.mark.parametrize("trans", "NT")
@pytest.mark.parametrize("n,nrhs", [(1, 1), (17, 3), (64, 16), (150, 2)])
def test_getrs_gesv_nopiv_batched_bitexact(queue, oracle, trans, n, nrhs):
    import magma_amd as ma
    import torch
    L = ma.load()
    rng = np.random.default_rng(n + nrhs)
    count = 4
    A = _dominant(rng, count, n, n)
    B = rng.uniform(-1, 1, (count, nrhs, n))
    dA = ma.DeviceBatch(A, ld=n + 1)
    info = torch.zeros(count, dtype=torch.int32, device="cuda")
    assert L.magma_dgetrf_nopiv_batched(n, n, dA.ptr, dA.ld, info.data_ptr(), count, queue.handle) == 0
    dB = ma.DeviceBatch(B, ld=n + 2)
    et = ma.MagmaNoTrans if trans == "N" else ma.MagmaTrans
    print(f"getrs_nopiv {trans} n={n} nrhs={nrhs}", flush=True)
    assert L.magma_dgetrs_nopiv_batched(et, n, nrhs, dA.ptr, dA.ld, dB.ptr, dB.ld, info.data_ptr(), count,
                                        queue.handle) == 0
    LU = dA.numpy(); X = dB.numpy()
    ident = np.arange(1, n + 1, dtype=np.int32)
    for s in range(count):
        assert same(X[s], oracle.dgetrs(LU[s], ident, B[s], trans.encode())), s
    if trans == "N":
        dA2 = ma.DeviceBatch(A, ld=n + 1)
        dB2 = ma.DeviceBatch(B, ld=n + 2)
        assert L.magma_dgesv_nopiv_batched(n, nrhs, dA2.ptr, dA2.ld, dB2.ptr, dB2.ld, info.data_ptr(), count,
                                           queue.handle) == 0
        assert same(dA2.numpy(), LU) and same(dB2.numpy(), X)
        for s in range(count):
            assert oracle.solve_error(A[s], X[s], B[s]) < 30


def test_nopiv_argument_errors(queue):
    import magma_amd as ma
    L = ma.load()
    q = queue.handle
    assert L.magma_dgetrf_nopiv_batched(-1, 4, 0, 4, 0, 1, q) == -1
    assert L.magma_dgetrf_nopiv_batched(4, -1, 0, 4, 0, 1, q) == -2
    assert L.magma_dgetrf_nopiv_batched(4, 4, 0, 3, 0, 1, q) == -4
    assert L.magma_dgetrs_nopiv_batched(ma.MagmaNoTrans, 4, 1, 0, 3, 0, 4, 0, 1, q) == -5
    assert L.magma_dgesv_nopiv_batched(4, 1, 0, 4, 0, 3, 0, 1, q) == -6


@pytest.mark.parametrize("n,nrhs", [(1, 1), (2, 3), (3, 1), (5, 2), (6, 1), (8, 4), (31, 2), (64, 3), (100, 1),
                                    (255, 2), (256, 1)])
def test_gesv_rbt_batched_bitexact(queue, oracle, n, nrhs):
    """Butterfly transform, no-pivot LU and solves: A and X bitwise the
    oracle's (same weights, same operation order); X meets the RBT tester's
    bound ||B - A X|| / (n ||A|| ||X|| eps) < 30 (testing/testing_zgesv_rbt.cpp:101-103)."""
    import magma_amd as ma
    import torch
    L = ma.load()
    rng = np.random.default_rng(n * 7 + nrhs)
    count = 5
    A = rng.uniform(-1, 1, (count, n, n))
    B = rng.uniform(-1, 1, (count, nrhs, n))
    dA = ma.DeviceBatch(A, ld=n + 3)
    dB = ma.DeviceBatch(B, ld=n + 1)
    info = torch.full((count,), -4, dtype=torch.int32, device="cuda")
    print(f"gesv_rbt n={n} nrhs={nrhs}", flush=True)
    assert L.magma_dgesv_rbt_batched(n, nrhs, dA.ptr, dA.ld, dB.ptr, dB.ld, info.data_ptr(), count,
                                     queue.handle) == 0
    Ag = dA.numpy(); X = dB.numpy(); inf = info.cpu().numpy()
    for s in range(count):
        Ao, Xo, info_o = oracle.dgesv_rbt(A[s], B[s])
        assert inf[s] == info_o == 0
        assert same(Ag[s], Ao), s
        assert same(X[s], Xo), s
        assert oracle.solve_error(A[s], X[s], B[s]) < 30, s


def test_gesv_rbt_argument_errors(queue):
    import magma_amd as ma
    L = ma.load()
    q = queue.handle
    assert L.magma_dgesv_rbt_batched(-1, 1, 0, 1, 0, 1, 0, 1, q) == -1
    assert L.magma_dgesv_rbt_batched(4, -1, 0, 4, 0, 4, 0, 1, q) == -2
    assert L.magma_dgesv_rbt_batched(4, 1, 0, 3, 0, 4, 0, 1, q) == -4
    assert L.magma_dgesv_rbt_batched(4, 1, 0, 4, 0, 3, 0, 1, q) == -6


# -------------------------------------------------------------------- QR ----
@pytest.mark.parametrize("m,n", [(1, 1), (2, 1), (5, 5), (32, 32), (64, 64), (100, 60), (60, 100), (256, 256),
                                 (300, 17)])
def test_geqrf_batched(queue, oracle, m, n):
    """The tester's bar (testing/testing_zgeqrf_batched.cpp:34-82, 116):
    ||R - Q'A||_1 / (n ||A||_1) < 30 eps and ||I - Q'Q||_1 / n < 30 eps, and
    R, tau equal to the oracle's (netlib dgeqr2) to rounding.  Not bitwise:
    the column norm is a two-pass scaled sum here, dnrm2's running scaled sum
    of squares in the oracle."""
    import magma_amd as ma
    import torch
    L = ma.load()
    rng = np.random.default_rng(m * 1000 + n)
    count = 5
    A = rng.uniform(-1, 1, (count, n, m))
    A[count - 1][0, :] = 0.0                          # a zero first column: tau[0] = 0
    dA = ma.DeviceBatch(A, ld=m + 2)
    k = min(m, n)
    tau = ma.DeviceBatch(np.full((count, 1, k), 9.0))
    info = torch.full((count,), -1, dtype=torch.int32, device="cuda")
    print(f"geqrf m={m} n={n}", flush=True)
    assert L.magma_dgeqrf_batched(m, n, dA.ptr, dA.ld, tau.ptr, info.data_ptr(), count, queue.handle) == 0
    F = dA.numpy(); T = tau.numpy()[:, 0, :]
    assert np.all(info.cpu().numpy() == 0)
    for s in range(count):
        e1, e2 = oracle.qr_errors(A[s], F[s], T[s], m, n)
        assert e1 < 30 * 2.0 ** -53 and e2 < 30 * 2.0 ** -53, (s, e1, e2)
        Fo, to = oracle.dgeqrf(A[s])
        scale = max(1.0, np.abs(Fo).max())
        assert np.abs(np.triu(F[s].T[:k, :]) - np.triu(Fo.T[:k, :])).max() <= 1e-12 * scale * max(m, n), s
        assert np.abs(T[s] - to).max() <= 1e-12 * max(m, n), s
    assert T[count - 1][0] == 0.0


def test_geqrf_argument_errors(queue):
    import magma_amd as ma
    L = ma.load()
    q = queue.handle
    assert L.magma_dgeqrf_batched(-1, 4, 0, 4, 0, 0, 1, q) == -1
    assert L.magma_dgeqrf_batched(4, -1, 0, 4, 0, 0, 1, q) == -2
    assert L.magma_dgeqrf_batched(4, 4, 0, 3, 0, 0, 1, q) == -4
