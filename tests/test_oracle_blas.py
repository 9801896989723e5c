"""CPU: the oracle's batched-BLAS restatements (dtrsm, dtrsv, dgemv) pinned
against closed forms and a third-party BLAS.

- Integer-exact systems: a triangular A with small integer entries and
  diagonal in {1, -1, 2, -2}, an integer X, B = op(A) X (or X op(A)) formed
  exactly in int64.  Every intermediate of a substitution is then an exact
  binary fraction, so the solve must return alpha X bit for bit.  The triangle
  that must not be read, and the diagonal when diag = 'U', hold NaN, so a
  transposed operand, a wrong triangle or a stray index shows up as NaN or a
  wrong integer.
- scipy's bundled OpenBLAS (third-party, not the reference) on random
  well-conditioned systems, to the testers' tolerance: it rounds in its own
  order, so the comparison is not bitwise.
"""
import itertools

import numpy as np
import pytest

EPS = 2.0 ** -53


def tri_storage(k, uplo, diag, rng, ld=None):
    """(cols, ld) storage of an order-k triangular matrix; returns (storage,
    A_math) with A_math the matrix the routine must see (unit diagonal when
    diag = 'U')."""
    ld = k if ld is None else ld
    M = rng.integers(-3, 4, (k, k)).astype(np.float64)
    d = rng.choice([1.0, -1.0, 2.0, -2.0], k)
    np.fill_diagonal(M, d)
    mask = np.triu(np.ones((k, k), bool)) if uplo == b"U" else np.tril(np.ones((k, k), bool))
    A_math = np.where(mask, M, 0.0)
    if diag == b"U":
        np.fill_diagonal(A_math, 1.0)
    S = np.full((k, ld), np.nan)
    stored = np.where(mask, M, np.nan)
    if diag == b"U":
        np.fill_diagonal(stored, np.nan)
    S[:, :k] = stored.T                       # storage[j, i] = A(i, j)
    return S, A_math


def op(A, trans):
    return A if trans == b"N" else A.T


CASES = list(itertools.product([b"L", b"R"], [b"U", b"L"], [b"N", b"T"], [b"N", b"U"]))


@pytest.mark.parametrize("side,uplo,trans,diag", CASES)
@pytest.mark.parametrize("alpha", [1.0, 2.0])
def test_trsm_integer_exact(oracle, side, uplo, trans, diag, alpha):
    rng = np.random.default_rng(CASES.index((side, uplo, trans, diag)))
    m, n = 7, 5
    k = m if side == b"L" else n
    S, Am = tri_storage(k, uplo, diag, rng, ld=k + 2)
    X = rng.integers(-5, 6, (m, n)).astype(np.float64)
    B = op(Am, trans) @ X if side == b"L" else X @ op(Am, trans)
    Bs = np.zeros((n, m + 1))
    Bs[:, :m] = B.T
    got = oracle.dtrsm(side, uplo, trans, diag, m, n, alpha, S, Bs)
    assert np.array_equal(got[:, :m].T, alpha * X)
    assert np.all(got[:, m] == 0.0)           # rows past m untouched


@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product([b"U", b"L"], [b"N", b"T"], [b"N", b"U"])))
@pytest.mark.parametrize("incx", [1, 3])
def test_trsv_integer_exact(oracle, uplo, trans, diag, incx):
    rng = np.random.default_rng(7)
    n = 9
    S, Am = tri_storage(n, uplo, diag, rng)
    x = rng.integers(-5, 6, n).astype(np.float64)
    b = op(Am, trans) @ x
    xs = np.full(n * incx, -7.0)
    xs[::incx] = b
    got = oracle.dtrsv(uplo, trans, diag, n, S, xs, incx)
    assert np.array_equal(got[::incx], x)
    if incx > 1:
        assert np.all(np.delete(got, np.arange(0, n * incx, incx)) == -7.0)


def test_trsm_alpha_zero_and_empty(oracle):
    S = np.full((3, 3), np.nan)
    B = np.ones((2, 3))
    assert np.array_equal(oracle.dtrsm(b"L", b"U", b"N", b"N", 3, 2, 0.0, S, B), np.zeros((2, 3)))
    assert np.array_equal(oracle.dtrsm(b"L", b"U", b"N", b"N", 0, 2, 1.0, S, B), B)


@pytest.mark.parametrize("trans", [b"N", b"T"])
@pytest.mark.parametrize("incx,incy", [(1, 1), (2, 3)])
def test_gemv_integer_exact(oracle, trans, incx, incy):
    rng = np.random.default_rng(3)
    m, n = 6, 9
    A = rng.integers(-4, 5, (m, n))
    lx, ly = (n, m) if trans == b"N" else (m, n)
    x = rng.integers(-4, 5, lx)
    y = rng.integers(-4, 5, ly)
    want = 3 * (op(A, trans) @ x) - 2 * y
    S = np.zeros((n, m + 1)); S[:, :m] = A.T
    xs = np.full(lx * incx, 5.0); xs[::incx] = x
    ys = np.full(ly * incy, 5.0); ys[::incy] = y
    got = oracle.dgemv(trans, m, n, 3.0, S, xs, -2.0, ys, incx, incy)
    assert np.array_equal(got[::incy], want.astype(np.float64))
    assert np.all(np.delete(got, np.arange(0, ly * incy, incy)) == 5.0)


def test_gemv_beta_zero_clears_nan_and_quick_return(oracle):
    A = np.ones((3, 2)); x = np.ones(3); y = np.full(2, np.nan)
    got = oracle.dgemv(b"N", 2, 3, 1.0, A, x, 0.0, y)
    assert np.array_equal(got, [3.0, 3.0])
    y2 = np.array([np.nan, 4.0])
    assert np.array_equal(oracle.dgemv(b"N", 2, 3, 0.0, A, x, 1.0, y2), y2, equal_nan=True)


def _well_conditioned(k, uplo, rng):
    M = rng.uniform(-1, 1, (k, k)) / k
    np.fill_diagonal(M, rng.uniform(1, 2, k) * rng.choice([-1, 1], k))
    return np.triu(M) if uplo == b"U" else np.tril(M)


@pytest.mark.parametrize("side,uplo,trans,diag", CASES)
def test_trsm_matches_openblas(oracle, side, uplo, trans, diag):
    from scipy.linalg import blas
    rng = np.random.default_rng(11)
    m, n = 33, 17
    k = m if side == b"L" else n
    Am = _well_conditioned(k, uplo, rng)
    B = rng.uniform(-1, 1, (m, n))
    ref = blas.dtrsm(0.7, Am, B, side=int(side == b"R"), lower=int(uplo == b"L"),
                     trans_a=int(trans == b"T"), diag=int(diag == b"U"))
    got = oracle.dtrsm(side, uplo, trans, diag, m, n, 0.7, np.ascontiguousarray(Am.T),
                       np.ascontiguousarray(B.T)).T
    assert np.max(np.abs(got - ref)) <= 30 * EPS * k * np.max(np.abs(ref))


@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product([b"U", b"L"], [b"N", b"T"], [b"N", b"U"])))
def test_trsv_matches_openblas(oracle, uplo, trans, diag):
    from scipy.linalg import blas
    rng = np.random.default_rng(12)
    n = 41
    Am = _well_conditioned(n, uplo, rng)
    b = rng.uniform(-1, 1, n)
    ref = blas.dtrsv(Am, b, lower=int(uplo == b"L"), trans=int(trans == b"T"), diag=int(diag == b"U"))
    got = oracle.dtrsv(uplo, trans, diag, n, np.ascontiguousarray(Am.T), b)
    assert np.max(np.abs(got - ref)) <= 30 * EPS * n * np.max(np.abs(ref))


@pytest.mark.parametrize("trans", [b"N", b"T"])
def test_gemv_matches_openblas(oracle, trans):
    from scipy.linalg import blas
    rng = np.random.default_rng(13)
    m, n = 37, 23
    A = rng.uniform(-1, 1, (m, n))
    x = rng.uniform(-1, 1, n if trans == b"N" else m)
    y = rng.uniform(-1, 1, m if trans == b"N" else n)
    ref = blas.dgemv(1.3, A, x, beta=-0.4, y=y.copy(), trans=int(trans == b"T"))
    got = oracle.dgemv(trans, m, n, 1.3, np.ascontiguousarray(A.T), x, -0.4, y)
    k = n if trans == b"N" else m
    # the gemv tester's bound (testing/testing_zgemv_batched.cpp:173-183)
    bound = 3 * EPS * (np.sqrt(k + 2) * np.linalg.norm(A) * np.linalg.norm(x) + 2 * np.linalg.norm(y))
    assert np.linalg.norm(got - ref) <= bound


# ------------------------------------------------------------ LU, no pivots --
@pytest.mark.parametrize("m,n", [(6, 6), (9, 5), (5, 9), (1, 1)])
def test_getrf_nopiv_integer_exact(oracle, m, n):
    """A = L U with L unit lower and U upper of small integers, diagonal of U
    in {1, -1, 2, -2}: the factorization returns L and U exactly."""
    rng = np.random.default_rng(m * 10 + n)
    k = min(m, n)
    Lm = np.tril(rng.integers(-3, 4, (m, k)).astype(np.float64), -1)
    Lm[np.arange(k), np.arange(k)] = 1.0
    Um = np.triu(rng.integers(-3, 4, (k, n)).astype(np.float64))
    Um[np.arange(k), np.arange(k)] = rng.choice([1.0, -1.0, 2.0, -2.0], k)
    A = Lm @ Um
    LU, info = oracle.dgetrf_nopiv(np.ascontiguousarray(A.T))
    got = LU.T
    assert info == 0
    assert np.array_equal(np.tril(got[:, :k], -1), np.tril(Lm, -1)[:, :k])
    assert np.array_equal(np.triu(got[:k, :]), Um)


def test_getrf_nopiv_equals_pivoted_when_no_swap(oracle):
    """Column-diagonally dominant matrices: partial pivoting never swaps, so
    the no-pivot factorization is the pivoted one bit for bit."""
    rng = np.random.default_rng(3)
    n = 40
    A = rng.uniform(-1, 1, (n, n))
    A[np.arange(n), np.arange(n)] = np.abs(A).sum(axis=0) + 1.0
    S = np.ascontiguousarray(A.T)
    LUp, ipiv, infop = oracle.dgetrf(S)
    assert np.array_equal(ipiv, np.arange(1, n + 1))
    LUn, info = oracle.dgetrf_nopiv(S)
    assert info == infop == 0 and np.array_equal(LUn, LUp)


def test_getrf_nopiv_zero_pivot_info(oracle):
    A = np.array([[0.0, 1.0], [1.0, 1.0]])          # A(0,0) = 0
    LU, info = oracle.dgetrf_nopiv(A.T.copy())
    assert info == 1
    B = np.array([[2.0, 1.0, 0.0], [4.0, 2.0, 1.0], [1.0, 3.0, 5.0]])   # second pivot 2 - 2 = 0
    LU, info = oracle.dgetrf_nopiv(B.T.copy())
    assert info == 2


# ------------------------------------------------------- butterfly solve ----
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 31, 64, 100])
def test_rbt_transform_is_u_a_v(oracle, n):
    """The transformed matrix factors as L U = U' A V and the solution solves
    A X = B, with U' and V built here as explicit products of butterfly
    matrices (each pairs x with x + N/2 and leaves an odd leftover alone)."""
    rng = np.random.default_rng(n)
    A = rng.uniform(-1, 1, (n, n)) + n * np.eye(n)        # dominant: no-pivot LU is accurate
    B = rng.uniform(-1, 1, (2, n))
    u, v = oracle.rbt_butterflies(n)

    def fly(N, w, off, size):
        M = np.eye(size)
        h = N // 2
        for x in range(h):
            i, j = off + x, off + x + h
            M[i, :] = 0; M[j, :] = 0
            M[i, i] = w[x]; M[i, j] = w[x]
            M[j, i] = w[x + h]; M[j, j] = -w[x + h]
        return M

    h = n // 2
    L1 = fly(h, u[n:], 0, n) @ fly(h, u[n + h:], h, n)
    R1 = fly(h, v[n:], 0, n) @ fly(h, v[n + h:], h, n)
    UT = fly(n, u, 0, n) @ L1                   # level 1 first, then level 2
    V = R1.T @ fly(n, v, 0, n).T
    Ao, Xo, info = oracle.dgesv_rbt(np.ascontiguousarray(A.T), B)
    LU = Ao.T
    Lf = np.tril(LU, -1) + np.eye(n); Uf = np.triu(LU)
    T = UT @ A @ V
    assert info == 0
    assert np.abs(Lf @ Uf - T).max() <= 1e-12 * max(1.0, np.abs(T).max())
    assert oracle.solve_error(np.ascontiguousarray(A.T), Xo, B) < 30


def test_rbt_butterflies_distribution(oracle):
    u, v = oracle.rbt_butterflies(500)
    for w in (u, v):
        assert w.shape == (1000,)
        assert np.all((w > np.exp(-0.05)) & (w < np.exp(0.05)))
    assert not np.array_equal(u, v)


# ------------------------------------------------------------------ QR ------
@pytest.mark.parametrize("m,n", [(1, 1), (5, 5), (40, 40), (60, 25), (25, 60), (128, 128)])
def test_geqrf_matches_lapack(oracle, m, n):
    """R and tau against LAPACK's dgeqrf (scipy's OpenBLAS, third-party) to
    rounding, and the tester's two residuals below 30 eps."""
    from scipy.linalg import qr
    rng = np.random.default_rng(m * 100 + n)
    A = rng.uniform(-1, 1, (m, n))
    (qr_raw, tau_ref), _ = qr(A, mode="raw")
    F, tau = oracle.dgeqrf(np.ascontiguousarray(A.T))
    k = min(m, n)
    assert np.allclose(np.triu(F.T[:k, :]), np.triu(qr_raw[:k, :]), rtol=0, atol=1e-13 * max(m, n))
    assert np.allclose(tau, tau_ref, rtol=0, atol=1e-13 * max(m, n))
    e1, e2 = oracle.qr_errors(np.ascontiguousarray(A.T), F, tau, m, n)
    assert e1 < 30 * EPS and e2 < 30 * EPS


def test_geqrf_closed_forms(oracle):
    # (3, 4)': beta = -5, tau = 8/5, v = (1, 1/2): all exact
    F, tau = oracle.dgeqrf(np.array([[3.0, 4.0]]))
    assert F[0, 0] == -5.0 and F[0, 1] == 0.5 and tau[0] == 1.6
    # upper triangular: nothing to annihilate, tau = 0 and R = A
    U = np.triu(np.arange(1.0, 17.0).reshape(4, 4))
    F, tau = oracle.dgeqrf(np.ascontiguousarray(U.T))
    assert np.all(tau == 0.0) and np.array_equal(F, U.T)
