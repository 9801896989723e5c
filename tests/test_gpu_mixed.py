"""GPU: magma_dsgesv_batched / magma_dsposv_batched (FP32 factorization with
split-FP16 MFMA updates, FP64 refinement, FP64 fallback).

Bar: every solution has backward error ||b - Ax||_inf / (||A||_inf ||x||_inf)
<= n * eps64 (oracle-computed, orc_solve_error <= 1); matrices within the
refinement's reach (cond <= 1e6) converge without fallback and leave A and B
unchanged; singular, NaN and very ill-conditioned matrices take the FP64 path
with dsgesv's iter codes and then meet the plain solver's bar (< 30), with A
holding the FP64 factors bit-exact as the plain path produces them.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_mixed(queue, A, B, kind="L"):
    import magma_amd as ma
    import torch
    count, n, _ = A.shape
    nrhs = B.shape[1]
    dA = ma.DeviceBatch(A)
    dB = ma.DeviceBatch(B)
    dX = ma.DeviceBatch(np.zeros_like(B))
    ipiv = ma.IntBatch(count, n)
    info = torch.zeros(count, dtype=torch.int32, device="cuda")
    it = torch.full((count,), 99, dtype=torch.int32, device="cuda")
    L = ma.load()
    if kind == "L":
        print(f"dsgesv count={count} n={n} nrhs={nrhs}", flush=True)
        ret = L.magma_dsgesv_batched(n, nrhs, dA.ptr, dA.ld, ipiv.ptr, dB.ptr, dB.ld, dX.ptr, dX.ld,
                                     info.data_ptr(), it.data_ptr(), count, queue.handle)
    else:
        print(f"dsposv count={count} n={n} nrhs={nrhs}", flush=True)
        ret = L.magma_dsposv_batched(ma.MagmaLower, n, nrhs, dA.ptr, dA.ld, dB.ptr, dB.ld, dX.ptr, dX.ld,
                                     info.data_ptr(), it.data_ptr(), count, queue.handle)
    assert ret == 0
    return dA.numpy(), dB.numpy(), dX.numpy(), ipiv.numpy(), info.cpu().numpy(), it.cpu().numpy()


def spd_cond_batch(seed, count, n, kmax=1e6):
    """SPD matrices U diag(s) U' with cond log-uniform in [1e2, kmax]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        U, _ = np.linalg.qr(rng.standard_normal((n, n)))
        kappa = 10.0 ** rng.uniform(2.0, np.log10(kmax))
        s = kappa ** (-np.arange(n) / max(n - 1, 1))
        S = (U * s) @ U.T
        out.append(np.ascontiguousarray((S + S.T) / 2))
    return np.stack(out)


@pytest.mark.parametrize("n,nrhs,count", [(256, 1, 48), (100, 2, 40), (33, 1, 40), (300, 1, 12), (1, 1, 4),
                                          (64, 5, 20), (512, 2, 6), (200, 16, 8), (130, 21, 6), (17, 3, 30)])
def test_dsgesv_cond_matrices(queue, oracle, n, nrhs, count):
    from magma_amd import gen
    A, kappa = gen.cond_batch(3, count, n) if n > 1 else (np.full((count, 1, 1), 0.5), np.ones(count))
    B = gen.rhs_batch(3, count, n, nrhs)
    Ag, Bg, X, ipiv, info, it = run_mixed(queue, A, B)
    assert np.all(info == 0)
    assert np.all(it >= 0), f"fallbacks: {np.sum(it < 0)} iter={it}"
    assert np.all(it <= 10)
    assert np.array_equal(Ag, A) and np.array_equal(Bg, B)       # untouched when refinement succeeds
    # the FP32 factorization's pivots are LAPACK-form: ipiv[i] in [i+1, n]
    idx = np.arange(1, n + 1)
    assert np.all((ipiv >= idx) & (ipiv <= n))
    for s in range(count):
        be = oracle.solve_error(A[s], X[s], B[s])                 # ||r|| / (n ||A|| ||x|| eps)
        assert be <= 1.0, (s, be, kappa[s], it[s])


def test_dsgesv_pivots_match_lapack_sgetrf_on_separated_columns(queue, oracle):
    """With one dominant entry per column the pivot choice has no near ties,
    so the FP32 factorization's interchanges are LAPACK's."""
    rng = np.random.default_rng(5)
    n, count = 96, 10
    mats = []
    for _ in range(count):
        perm = rng.permutation(n)
        M = rng.uniform(-1, 1, (n, n)) * 1e-3
        M[perm, np.arange(n)] = rng.uniform(1, 2, n) * rng.choice([-1, 1], n)
        mats.append(np.ascontiguousarray(M.T))
    A = np.stack(mats)
    B = np.ones((count, 1, n))
    Ag, Bg, X, ipiv, info, it = run_mixed(queue, A, B)
    assert np.all(it >= 0)
    for s in range(count):
        _, ipo, _ = oracle.dgetrf(A[s])
        assert np.array_equal(ipiv[s], ipo), s


def test_dsgesv_fallbacks(queue, oracle):
    from magma_amd import gen
    n, count = 96, 8
    A, _ = gen.cond_batch(5, count, n)
    A = A.copy()
    A[1] = 0.0                                                    # zero matrix: singular (iter -3, info 1)
    A[2][5, :] = A[2][4, :]                                       # singular: two equal columns
    A[3][7, 9] = np.nan                                           # cannot be scaled (iter -2)
    # condition 1e13: out of the refinement's reach (iter -11), FP64 still solves it
    U, _ = np.linalg.qr(np.random.default_rng(0).standard_normal((n, n)))
    V, _ = np.linalg.qr(np.random.default_rng(1).standard_normal((n, n)))
    s = 10.0 ** (-13 * np.arange(n) / (n - 1))
    A[4] = np.ascontiguousarray(((U * s) @ V.T).T)
    B = gen.rhs_batch(5, count, n, 1)
    Ag, Bg, X, ipiv, info, it = run_mixed(queue, A, B)
    assert it[0] >= 0 and it[5] >= 0
    assert it[1] == -3 and info[1] == 1
    assert it[2] < 0                                              # singular: FP32 pivot may be tiny, not zero
    assert it[3] == -2
    assert it[4] == -(10 + 1)
    # the fallback factor is LAPACK's: A holds the FP64 LU and ipiv its pivots
    LUo, ipo, info_o = oracle.dgetrf(A[4])
    assert np.array_equal(ipiv[4], ipo) and np.array_equal(Ag[4], LUo)
    assert oracle.solve_error(A[4], X[4], B[4]) < 30
    Xo = oracle.dgetrs(LUo, ipo, B[4])
    assert np.array_equal(X[4], Xo)


def test_dsgesv_large_n_falls_back(queue, oracle):
    from magma_amd import gen
    n, count = 600, 2
    A = gen.general_batch(8, count, n)
    B = gen.rhs_batch(8, count, n, 1)
    Ag, Bg, X, ipiv, info, it = run_mixed(queue, A, B)
    assert np.all(it == -1)
    for s in range(count):
        assert oracle.solve_error(A[s], X[s], B[s]) < 30


@pytest.mark.parametrize("n,nrhs,count", [(1, 1, 4), (20, 3, 30), (64, 4, 20), (100, 1, 20), (256, 2, 12),
                                          (300, 5, 6), (512, 1, 4)])
def test_dsposv_spd(queue, oracle, n, nrhs, count):
    from magma_amd import gen
    A = gen.spd_batch(4, count, n)
    B = gen.rhs_batch(4, count, n, nrhs)
    Ag, Bg, X, ipiv, info, it = run_mixed(queue, A, B, kind="C")
    assert np.all(info == 0)
    assert np.all((it >= 0) & (it <= 10)), it
    assert np.array_equal(Ag, A) and np.array_equal(Bg, B)
    for s in range(count):
        assert oracle.solve_error(A[s], X[s], B[s]) <= 1.0, s


@pytest.mark.parametrize("n", [48, 200])
def test_dsposv_conditioned(queue, oracle, n):
    from magma_amd import gen
    A = spd_cond_batch(9, 10, n)
    B = gen.rhs_batch(9, 10, n, 2)
    Ag, Bg, X, ipiv, info, it = run_mixed(queue, A, B, kind="C")
    assert np.all(info == 0)
    assert np.all(it >= 0), it
    for s in range(10):
        assert oracle.solve_error(A[s], X[s], B[s]) <= 1.0, s


def test_dsposv_not_spd_falls_back(queue, oracle):
    from magma_amd import gen
    n, count = 70, 4
    A = gen.spd_batch(6, count, n).copy()
    A[1][10, 10] = -5.0                   # negative diagonal: not SPD before any update
    A[2] = -A[2]                          # negative definite
    A[3][20, 30] = A[3][30, 20] = 1e4     # indefinite after elimination
    B = gen.rhs_batch(6, count, n, 1)
    Ag, Bg, X, ipiv, info, it = run_mixed(queue, A, B, kind="C")
    assert it[0] >= 0 and info[0] == 0
    for s in (1, 2, 3):
        assert it[s] < 0, (s, it[s])
        Lo, info_o = oracle.dpotrf(A[s])
        assert info[s] == info_o and info_o > 0, (s, info[s], info_o)


def test_dsgesv_argument_errors(queue):
    import magma_amd as ma
    L = ma.load()
    q = queue.handle
    assert L.magma_dsgesv_batched(-1, 1, 0, 1, 0, 0, 1, 0, 1, 0, 0, 1, q) == -1
    assert L.magma_dsgesv_batched(4, -1, 0, 4, 0, 0, 4, 0, 4, 0, 0, 1, q) == -2
    assert L.magma_dsgesv_batched(4, 1, 0, 3, 0, 0, 4, 0, 4, 0, 0, 1, q) == -4
    assert L.magma_dsgesv_batched(4, 1, 0, 4, 0, 0, 3, 0, 4, 0, 0, 1, q) == -7
    assert L.magma_dsgesv_batched(4, 1, 0, 4, 0, 0, 4, 0, 3, 0, 0, 1, q) == -9
    # dsposv keeps magma_dsposv_gpu's codes (n, nrhs, ldda numbered without uplo)
    assert L.magma_dsposv_batched(0, 4, 1, 0, 4, 0, 4, 0, 4, 0, 0, 1, q) == -1           # no such uplo
    assert L.magma_dsposv_batched(ma.MagmaLower, -1, 1, 0, 4, 0, 4, 0, 4, 0, 0, 1, q) == -1
    assert L.magma_dsposv_batched(ma.MagmaLower, 4, -1, 0, 4, 0, 4, 0, 4, 0, 0, 1, q) == -2
    assert L.magma_dsposv_batched(ma.MagmaLower, 4, 1, 0, 3, 0, 4, 0, 4, 0, 0, 1, q) == -4
    assert L.magma_dsposv_batched(ma.MagmaLower, 4, 1, 0, 4, 0, 3, 0, 4, 0, 0, 1, q) == -7


@pytest.mark.parametrize("n,nrhs", [(0, 1), (4, 0), (0, 0)])
def test_quick_return_sets_iter_and_info(queue, n, nrhs):
    """n = 0 or nrhs = 0: nothing is factored or solved, and every matrix
    gets iter = 0, info = 0, as magma_dsgesv_gpu / magma_dsposv_gpu leave
    them (src/zcgesv_gpu.cpp:174-193, src/zcposv_gpu.cpp:163-182)."""
    import magma_amd as ma
    L = ma.load()
    count = 5
    for kind in ("L", "C"):
        info = ma.dev_ints(np.full(count, 77))
        it = ma.dev_ints(np.full(count, 99))
        print(f"quick return kind={kind} n={n} nrhs={nrhs}", flush=True)
        if kind == "L":
            ret = L.magma_dsgesv_batched(n, nrhs, 0, 4, 0, 0, 4, 0, 4, info.data_ptr(), it.data_ptr(), count, queue.handle)
        else:
            ret = L.magma_dsposv_batched(ma.MagmaLower, n, nrhs, 0, 4, 0, 4, 0, 4, info.data_ptr(), it.data_ptr(),
                                         count, queue.handle)
        assert ret == 0
        # no queue.sync(): the call itself synchronizes before returning
        assert np.all(info.cpu().numpy()[:count] == 0) and np.all(it.cpu().numpy()[:count] == 0), kind


# ---------------------------------------------------------------- vbatched ---
def run_mixed_v(queue, mats, rhs, kind="L"):
    """magma_dsgesv_vbatched / magma_dsposv_vbatched on matrices of different
    orders; returns per-matrix A, X, ipiv, info, iter."""
    import magma_amd as ma
    count = len(mats)
    V = ma.VarBatch(mats)
    R = ma.VarBatch(rhs)
    Xv = ma.VarBatch([np.zeros_like(b) for b in rhs])
    sizes = [m.shape[0] for m in mats]
    ipiv = ma.VarInts([max(n, 1) for n in sizes])
    info = ma.dev_ints(np.full(count, 77))
    it = ma.dev_ints(np.full(count, 99))
    L = ma.load()
    print(f"{'dsgesv' if kind == 'L' else 'dsposv'}_vbatched count={count} sizes={sizes}", flush=True)
    if kind == "L":
        ret = L.magma_dsgesv_vbatched(V.d_cols.data_ptr(), R.d_cols.data_ptr(), V.ptr, V.d_ld.data_ptr(),
                                      ipiv.ptr, R.ptr, R.d_ld.data_ptr(), Xv.ptr, Xv.d_ld.data_ptr(),
                                      info.data_ptr(), it.data_ptr(), count, queue.handle)
    else:
        ret = L.magma_dsposv_vbatched(ma.MagmaLower, V.d_cols.data_ptr(), R.d_cols.data_ptr(), V.ptr,
                                      V.d_ld.data_ptr(), R.ptr, R.d_ld.data_ptr(), Xv.ptr, Xv.d_ld.data_ptr(),
                                      info.data_ptr(), it.data_ptr(), count, queue.handle)
    assert ret == 0
    return (V.matrices(), Xv.matrices(), [ipiv.numpy(i, n) for i, n in enumerate(sizes)],
            info.cpu().numpy()[:count], it.cpu().numpy()[:count])


def _cond_matrix(seed, n):
    from magma_amd import gen
    if n == 0:
        return np.zeros((0, 0))
    if n == 1:
        return np.full((1, 1), 0.5)
    return gen.cond_batch(seed, 1, n)[0][0]


def test_dsgesv_vbatched_matches_batched(queue, oracle):
    """Each matrix of a mixed-order batch gets the bits the fixed-size call
    gives it (same kernel, same arithmetic), including the FP64 fallbacks;
    orders above 512 take the FP64 path with iter = -1."""
    from magma_amd import gen
    sizes = [256, 1, 17, 33, 64, 65, 100, 128, 129, 200, 255, 300, 512, 0, 600, 8, 9, 96, 96]
    nrhs = [1, 2, 3, 5, 1, 2, 1, 4, 1, 1, 2, 3, 1, 1, 1, 2, 1, 1, 1]
    mats = [_cond_matrix(20 + i, n) for i, n in enumerate(sizes)]
    mats[17] = np.zeros((96, 96))                                     # singular: iter -3, info 1
    U, _ = np.linalg.qr(np.random.default_rng(0).standard_normal((96, 96)))
    W, _ = np.linalg.qr(np.random.default_rng(1).standard_normal((96, 96)))
    mats[18] = np.ascontiguousarray(((U * 10.0 ** (-13 * np.arange(96) / 95)) @ W.T).T)   # cond 1e13
    rhs = [gen.rhs_batch(40 + i, 1, n, r)[0] if n else np.zeros((r, 0)) for i, (n, r) in enumerate(zip(sizes, nrhs))]
    Ag, X, ipiv, info, it = run_mixed_v(queue, mats, rhs)
    for i, n in enumerate(sizes):
        if n == 0:
            assert it[i] == 0 and info[i] == 0
            continue
        if n > 512:
            assert it[i] == -1 and info[i] == 0, (i, it[i])
            assert oracle.solve_error(mats[i], X[i], rhs[i]) < 30, i
            continue
        if i == 17:
            assert it[i] == -3 and info[i] == 1
            continue
        if i == 18:
            assert it[i] == -11
            LUo, ipo, _ = oracle.dgetrf(mats[i])
            assert np.array_equal(ipiv[i], ipo) and np.array_equal(Ag[i], LUo)
            assert np.array_equal(X[i], oracle.dgetrs(LUo, ipo, rhs[i]))
            continue
        Ab, Bb, Xb, ipb, infob, itb = run_mixed(queue, mats[i][None], rhs[i][None])
        assert it[i] == itb[0] and 0 <= it[i] <= 10, (i, it[i], itb[0])
        assert info[i] == 0
        assert np.array_equal(X[i], Xb[0]), i
        assert np.array_equal(ipiv[i], ipb[0]), i
        assert np.array_equal(Ag[i], mats[i]), i                      # untouched when refinement succeeds
        assert oracle.solve_error(mats[i], X[i], rhs[i]) <= 1.0, i


def test_dsposv_vbatched_matches_batched(queue, oracle):
    from magma_amd import gen
    sizes = [1, 20, 64, 65, 100, 256, 300, 512, 0, 600, 70]
    nrhs = [1, 3, 4, 1, 2, 1, 5, 1, 1, 2, 1]
    mats = [gen.spd_batch(50 + i, 1, n)[0] if n else np.zeros((0, 0)) for i, n in enumerate(sizes)]
    mats[10] = mats[10].copy()
    mats[10][10, 10] = -5.0                                           # not SPD: FP64 path, info from potrf
    rhs = [gen.rhs_batch(60 + i, 1, n, r)[0] if n else np.zeros((r, 0)) for i, (n, r) in enumerate(zip(sizes, nrhs))]
    Ag, X, _, info, it = run_mixed_v(queue, mats, rhs, kind="C")
    for i, n in enumerate(sizes):
        if n == 0:
            assert it[i] == 0 and info[i] == 0
            continue
        if n > 512:
            assert it[i] == -1 and info[i] == 0
            assert oracle.solve_error(mats[i], X[i], rhs[i]) < 30, i
            continue
        if i == 10:
            _, info_o = oracle.dpotrf(mats[i])
            assert it[i] < 0 and info[i] == info_o and info_o > 0
            continue
        Ab, Bb, Xb, _, infob, itb = run_mixed(queue, mats[i][None], rhs[i][None], kind="C")
        assert it[i] == itb[0] and 0 <= it[i] <= 10, (i, it[i], itb[0])
        assert info[i] == 0
        assert np.array_equal(X[i], Xb[0]), i
        assert oracle.solve_error(mats[i], X[i], rhs[i]) <= 1.0, i


def test_mixed_vbatched_argument_errors(queue):
    import magma_amd as ma
    from magma_amd import lib as ma_lib
    L = ma.load()
    q = queue.handle
    ok = ma.dev_ints([4, 4, 4])
    neg = ma.dev_ints([4, -1, 4])
    small = ma.dev_ints([4, 3, 4])
    p = lambda t: t.data_ptr()

    def sg(n=ok, r=ok, la=ok, lb=ok, lx=ok, count=3):
        return L.magma_dsgesv_vbatched(p(n), p(r), 0, p(la), 0, 0, p(lb), 0, p(lx), 0, 0, count, q)

    assert sg(n=neg) == -1
    assert sg(r=neg) == -2
    assert sg(la=small) == -4
    assert sg(lb=small) == -7
    assert sg(lx=small) == -9
    assert sg(count=-1) == -12
    assert L.magma_dsposv_vbatched(0, p(ok), p(ok), 0, p(ok), 0, p(ok), 0, p(ok), 0, 0, 3, q) == -1      # no such uplo
    assert L.magma_dsposv_vbatched(ma.MagmaLower, p(neg), p(ok), 0, p(ok), 0, p(ok), 0, p(ok), 0, 0, 3, q) == -1
    assert L.magma_dsposv_vbatched(ma.MagmaLower, p(ok), p(ok), 0, p(small), 0, p(ok), 0, p(ok), 0, 0, 3, q) == -4
    big = ma.dev_ints([4, 9000, 4])
    assert sg(n=big, la=big, lb=big, lx=big) == ma_lib.MAGMA_ERR_NOT_SUPPORTED
