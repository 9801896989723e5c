"""Pins the CPU oracle (oracle/oracle.c) before anything is compared against it.

The reference holds no golden vectors for this path (its testers re-derive
parity at run time against whatever LAPACK is linked), so the oracle is pinned
by: closed-form factorizations (identity, reversal and cyclic permutations,
diagonal, zero columns), netlib idamax tie/NaN rules, the reciprocal scaling
and fma rounding contract, exact rational arithmetic on small integer
matrices (brute force), the reference testers' error formulas, and a
third-party LAPACK (scipy's OpenBLAS) on random inputs.
"""
import itertools
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg as sl
import scipy.linalg.lapack as sll

EPS = 2.0 ** -53


def cm(A):
    """row-major (n x n) numpy -> column-major storage (cols, rows)"""
    return np.ascontiguousarray(np.asarray(A, dtype=np.float64).T)


def rm(S):
    return np.ascontiguousarray(np.asarray(S).T)


def fma_exact(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


# ---------------------------------------------------------------- idamax ----
def test_idamax_first_maximum(oracle):
    assert oracle.idamax([1.0, -3.0, 3.0, 2.0]) == 1
    assert oracle.idamax([0.0, 0.0, 0.0]) == 0
    assert oracle.idamax([-5.0]) == 0


def test_idamax_nan_rules(oracle):
    nan = float("nan")
    assert oracle.idamax([nan, 1.0, 7.0]) == 0          # NaN first: selected
    assert oracle.idamax([1.0, nan, 7.0]) == 2          # later NaN never selected
    assert oracle.idamax([0.0, nan, nan]) == 0
    assert oracle.idamax([1.0, float("inf"), nan]) == 1


# ---------------------------------------------------------------- getrf -----
def test_getrf_identity(oracle):
    for n in (1, 2, 7, 33):
        LU, ipiv, info = oracle.dgetrf(cm(np.eye(n)))
        assert info == 0
        assert list(ipiv) == list(range(1, n + 1))
        assert np.array_equal(rm(LU), np.eye(n))


@pytest.mark.parametrize("n", [2, 3, 6, 9, 32])
def test_getrf_reversal_permutation(oracle, n):
    # column j of the reversal matrix has its 1 in row n-1-j: the first n/2
    # steps swap row j with row n-1-j, the rest are already in place.
    J = np.eye(n)[::-1]
    LU, ipiv, info = oracle.dgetrf(cm(J))
    expect = [n - j if j < n // 2 else j + 1 for j in range(n)]
    assert info == 0 and list(ipiv) == expect
    assert np.array_equal(rm(LU), np.eye(n))


@pytest.mark.parametrize("n", [2, 5, 8])
def test_getrf_cyclic_shift(oracle, n):
    # P e_j = e_{j+1 mod n}: column j has its 1 in row j+1 (mod n); step j
    # finds it in the row that the previous swaps moved down: ipiv[j] = j+2.
    P = np.zeros((n, n))
    for j in range(n):
        P[(j + 1) % n, j] = 1.0
    LU, ipiv, info = oracle.dgetrf(cm(P))
    assert info == 0
    assert list(ipiv) == [j + 2 for j in range(n - 1)] + [n]
    assert np.array_equal(rm(LU), np.eye(n))


def test_getrf_ties_pick_first(oracle):
    A = np.array([[1.0, 2.0, 0.0], [-1.0, 0.0, 1.0], [1.0, 1.0, 1.0]])
    LU, ipiv, info = oracle.dgetrf(cm(A))
    assert ipiv[0] == 1                                # |1| = |-1| = |1|: first row
    assert info == 0


def test_getrf_zero_column_continues(oracle):
    # column 1 becomes exactly zero below the diagonal -> info = 2, no swap,
    # later columns still factored
    A = np.array([[2.0, 4.0, 1.0], [1.0, 2.0, 3.0], [4.0, 8.0, 5.0]])
    LU, ipiv, info = oracle.dgetrf(cm(A))
    assert info == 2
    assert ipiv[0] == 3 and ipiv[1] == 2
    assert oracle.lu_error(cm(A), LU, ipiv) < 30
    D = np.diag([1.0, 0.0, 2.0])
    LU, ipiv, info = oracle.dgetrf(cm(D))
    assert info == 2 and list(ipiv) == [1, 2, 3] and np.array_equal(rm(LU), D)


def test_getrf_nan_not_selected(oracle):
    nan = float("nan")
    A = np.array([[1.0, 1.0], [nan, 1.0], [2.0, 0.0]])
    LU, ipiv, info = oracle.dgetrf(cm(A), m=3, n=2)
    assert ipiv[0] == 3


def test_getrf_reciprocal_scaling(oracle):
    # netlib dgetf2 scales the column by 1/pivot: 3 * (1/17) != 3/17 in binary64
    assert 3.0 * (1.0 / 17.0) != 3.0 / 17.0
    A = np.array([[17.0, 1.0], [3.0, 1.0]])
    LU, ipiv, info = oracle.dgetrf(cm(A))
    l = rm(LU)[1, 0]
    assert l == 3.0 * (1.0 / 17.0)
    # division below sfmin
    tiny = 2.0 ** -1060
    A = np.array([[tiny, 0.0], [tiny / 3.0, 1.0]])
    LU, ipiv, info = oracle.dgetrf(cm(A))
    assert rm(LU)[1, 0] == (tiny / 3.0) / tiny


def test_getrf_update_is_fused(oracle):
    rng = np.random.default_rng(7)
    hits = 0
    for _ in range(200):
        p, u, a, l0 = rng.uniform(1.0, 2.0), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-0.9, 0.9)
        A = np.array([[p, u], [l0, a]])
        LU, ipiv, info = oracle.dgetrf(cm(A))
        l = l0 * (1.0 / p)
        got = rm(LU)[1, 1]
        assert got == fma_exact(l, -u, a)
        hits += got != a - l * u
    assert hits > 0          # the fused and unfused roundings do differ on this sample


def test_getrf_empty_and_scalar(oracle):
    LU, ipiv, info = oracle.dgetrf(np.zeros((1, 1)), m=0, n=0)
    assert info == 0 and len(ipiv) == 0
    LU, ipiv, info = oracle.dgetrf(cm([[5.0]]))
    assert info == 0 and list(ipiv) == [1] and rm(LU)[0, 0] == 5.0
    LU, ipiv, info = oracle.dgetrf(cm([[0.0]]))
    assert info == 1 and list(ipiv) == [1]


def _exact_lu_pivots(A):
    """Partial pivoting in exact rational arithmetic; returns (ipiv, tie flags)."""
    n = len(A)
    M = [[Fraction(x) for x in row] for row in A]
    ipiv, ties = [], []
    for j in range(n):
        col = [abs(M[i][j]) for i in range(j, n)]
        best = max(col)
        p = j + col.index(best)
        ties.append(col.count(best) > 1)
        ipiv.append(p + 1)
        if best == 0:
            continue
        M[j], M[p] = M[p], M[j]
        for i in range(j + 1, n):
            f = M[i][j] / M[j][j]
            M[i][j] = f
            for c in range(j + 1, n):
                M[i][c] -= f * M[j][c]
    return ipiv, ties


def test_getrf_brute_force_small(oracle):
    """All sign patterns of a fixed magnitude pattern (n=3: 2^9 matrices) and
    random small-integer 4x4 matrices: pivots equal exact-arithmetic partial
    pivoting wherever the exact column maximum is unique."""
    mags = np.array([[3.0, 1.0, 2.0], [1.0, 4.0, 1.0], [2.0, 2.0, 5.0]])
    checked = 0
    for signs in itertools.product((1.0, -1.0), repeat=9):
        A = mags * np.array(signs).reshape(3, 3)
        _, ipiv, _ = oracle.dgetrf(cm(A))
        ex, ties = _exact_lu_pivots(A.tolist())
        for j in range(3):
            if ties[j]:
                break
            assert ipiv[j] == ex[j]
            checked += 1
    rng = np.random.default_rng(11)
    for _ in range(300):
        A = rng.integers(-4, 5, size=(4, 4)).astype(float)
        _, ipiv, info = oracle.dgetrf(cm(A))
        ex, ties = _exact_lu_pivots(A.tolist())
        for j in range(4):
            if ties[j]:
                break
            assert ipiv[j] == ex[j]
            checked += 1
    assert checked > 1000


@pytest.mark.parametrize("n", [1, 5, 32, 64, 100, 256])
def test_getrf_matches_lapack(oracle, n):
    rng = np.random.default_rng(n)
    for t in range(3):
        A = rng.uniform(-1, 1, (n, n))
        LU, ipiv, info = oracle.dgetrf(cm(A))
        lu_s, piv_s, info_s = sll.dgetrf(A)
        assert info == info_s
        assert np.array_equal(ipiv, piv_s + 1)
        assert np.allclose(rm(LU), lu_s, rtol=0, atol=1e-11 * n)
        assert oracle.lu_error(cm(A), LU, ipiv) < 30


def test_lu_error_detects_wrong_pivots(oracle):
    rng = np.random.default_rng(3)
    A = rng.uniform(-1, 1, (8, 8))
    LU, ipiv, info = oracle.dgetrf(cm(A))
    assert oracle.lu_error(cm(A), LU, ipiv) < 30
    bad = ipiv.copy()
    bad[0] = 8 if bad[0] != 8 else 7
    assert oracle.lu_error(cm(A), LU, bad) > 1e3


# ---------------------------------------------------------------- getrs -----
@pytest.mark.parametrize("trans", [b"N", b"T"])
@pytest.mark.parametrize("n,nrhs", [(1, 1), (7, 3), (32, 1), (64, 4), (129, 2)])
def test_getrs_matches_lapack(oracle, n, nrhs, trans):
    rng = np.random.default_rng(100 + n)
    A = rng.uniform(-1, 1, (n, n))
    B = rng.uniform(-1, 1, (n, nrhs))
    LU, ipiv, info = oracle.dgetrf(cm(A))
    X = oracle.dgetrs(LU, ipiv, cm(B), trans=trans)
    lu_s, piv_s, _ = sll.dgetrf(A)
    Xs, _ = sll.dgetrs(lu_s, piv_s, B, trans=0 if trans == b"N" else 1)
    assert np.allclose(rm(X), Xs, rtol=1e-10, atol=1e-12)
    At = A if trans == b"N" else A.T
    assert oracle.solve_error(cm(At), X, cm(B)) < 30


def test_getrs_triangular_exact(oracle):
    # unit-lower L times upper U with exactly representable solve
    L = np.array([[1.0, 0, 0], [0.5, 1, 0], [0.25, 0.5, 1]])
    U = np.array([[2.0, 1, 1], [0, 4, 2], [0, 0, 8]])
    x = np.array([1.0, -2.0, 3.0])
    b = L @ U @ x
    LU = np.tril(L, -1) + U
    X = oracle.dgetrs(cm(LU), np.array([1, 2, 3], np.int32), cm(b.reshape(3, 1)))
    assert np.array_equal(rm(X)[:, 0], x)


def test_solve_error_detects_bad_solution(oracle):
    rng = np.random.default_rng(5)
    A = rng.uniform(-1, 1, (16, 16))
    B = rng.uniform(-1, 1, (16, 1))
    X = np.linalg.solve(A, B)
    assert oracle.solve_error(cm(A), cm(X), cm(B)) < 30
    X[3, 0] += 1e-8
    assert oracle.solve_error(cm(A), cm(X), cm(B)) > 1e3


# ---------------------------------------------------------------- potrf -----
def spd(rng, n):
    Bm = rng.standard_normal((n, n))
    return Bm @ Bm.T + n * np.eye(n)


@pytest.mark.parametrize("n", [1, 4, 33, 64, 200])
def test_potrf_matches_lapack(oracle, n):
    rng = np.random.default_rng(n)
    A = spd(rng, n)
    L, info = oracle.dpotrf(cm(A))
    Ls, info_s = sll.dpotrf(A, lower=1)
    assert info == info_s == 0
    assert np.allclose(np.tril(rm(L)), np.tril(Ls), rtol=1e-12, atol=1e-12)
    assert oracle.chol_error(cm(A), L) < 30
    # tester metric: ||L - L_lapack||_F / ||L_lapack||_F < 30 eps
    d = np.linalg.norm(np.tril(rm(L)) - np.tril(Ls)) / np.linalg.norm(np.tril(Ls))
    assert d < 30 * EPS


def test_potrf_info_closed_forms(oracle):
    _, info = oracle.dpotrf(cm(np.diag([1.0, 1.0, -1.0, 1.0])))
    assert info == 3
    _, info = oracle.dpotrf(cm(np.diag([4.0, 0.0, 1.0])))
    assert info == 2
    _, info = oracle.dpotrf(cm(np.array([[1.0, 1.0], [1.0, 1.0]])))
    assert info == 2                               # second leading minor singular
    _, info = oracle.dpotrf(cm(np.diag([1.0, float("nan"), 1.0])))
    assert info == 2
    # leading (info-1) block is final: compare it with LAPACK's
    rng = np.random.default_rng(2)
    A = spd(rng, 6)
    A[4, 4] = -50.0
    L, info = oracle.dpotrf(cm(A))
    Ls, info_s = sll.dpotrf(A, lower=1)
    assert info == info_s == 5
    k = info - 1
    assert np.allclose(np.tril(rm(L)[:k, :k]), np.tril(Ls[:k, :k]), rtol=1e-13)


def test_potrf_upper_mirror(oracle):
    rng = np.random.default_rng(9)
    A = spd(rng, 12)
    U, info = oracle.dpotrf(cm(A), uplo=b"U")
    L, info2 = oracle.dpotrf(cm(A), uplo=b"L")
    assert info == info2 == 0
    assert np.array_equal(np.triu(rm(U)), np.tril(rm(L)).T)


@pytest.mark.parametrize("uplo", [b"L", b"U"])
def test_potrs_matches_lapack(oracle, uplo):
    rng = np.random.default_rng(21)
    n, nrhs = 40, 4
    A = spd(rng, n)
    B = rng.uniform(-1, 1, (n, nrhs))
    F, info = oracle.dpotrf(cm(A), uplo=uplo)
    X = oracle.dpotrs(F, cm(B), uplo=uplo)
    assert np.allclose(rm(X), np.linalg.solve(A, B), rtol=1e-11, atol=1e-13)
    assert oracle.solve_error(cm(A), X, cm(B)) < 30


# ---------------------------------------------------------------- gemm ------
@pytest.mark.parametrize("ta,tb", [(b"N", b"N"), (b"N", b"T"), (b"T", b"N"), (b"T", b"T")])
def test_gemm_integer_exact(oracle, ta, tb):
    rng = np.random.default_rng(4)
    m, n, k = 7, 5, 9
    A = rng.integers(-8, 9, (m, k) if ta == b"N" else (k, m)).astype(float)
    B = rng.integers(-8, 9, (k, n) if tb == b"N" else (n, k)).astype(float)
    C = rng.integers(-8, 9, (m, n)).astype(float)
    opA = A if ta == b"N" else A.T
    opB = B if tb == b"N" else B.T
    got = oracle.dgemm(ta, tb, m, n, k, 2.0, cm(A), A.shape[0], cm(B), B.shape[0], -1.0, cm(C), m)
    assert np.array_equal(rm(got), 2.0 * opA @ opB - C)


# ------------------------------------------------- batched drivers / LAPACK --
def test_batched_dgesv_equals_single(oracle):
    from magma_amd import gen
    A = gen.general_batch(0, 50, 32)
    B = gen.rhs_batch(0, 50, 32, 1)
    LU, ipiv, X, info = oracle.batch_dgesv(A, B)
    for s in (0, 17, 49):
        LU1, ip1, inf1 = oracle.dgetrf(A[s])
        assert np.array_equal(LU1, LU[s]) and np.array_equal(ip1, ipiv[s]) and inf1 == info[s]
        X1 = oracle.dgetrs(LU1, ip1, B[s])
        assert np.array_equal(X1, X[s])


def test_batched_against_openblas(oracle):
    if oracle.lapack() is None:
        pytest.skip("no OpenBLAS with scipy_ symbols")
    from magma_amd import gen
    A = gen.general_batch(0, 200, 32)
    B = gen.rhs_batch(0, 200, 32, 1)
    LU, ipiv, X, info = oracle.batch_dgesv(A, B)
    LU2, ipiv2, X2, info2 = oracle.lapack_batch_dgesv(A, B)
    assert np.array_equal(ipiv, ipiv2) and np.array_equal(info, info2)
    assert np.allclose(X, X2, rtol=1e-9, atol=1e-12)
    S = gen.spd_batch(1, 50, 64)
    R = gen.rhs_batch(1, 50, 64, 4)
    L, Xs, infos = oracle.batch_dposv(S, R)
    L2, Xs2, infos2 = oracle.lapack_batch_dposv(S, R)
    assert np.array_equal(infos, infos2)
    assert np.allclose(Xs, Xs2, rtol=1e-10, atol=1e-13)
