"""GPU parity: variable-size (vbatched) LU / Cholesky / solves and the batched
and vbatched DGEMM, against the CPU oracle.

Bar: per matrix, ipiv and info bit-exact and factors bit-exact (same update
order as the oracle); GEMM against the reference tester's bound
||C - C_ref||_F / (sqrt(k+2) |A||B| + 2|C|) < 3 eps
(testing/testing_zgemm_vbatched.cpp:313-331) and bit-exact on integer inputs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -53


def same(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def sizes_cfg3(count, seed=2, lo=8, hi=256):
    from magma_amd import gen
    return gen.vbatch_sizes(seed, count, lo, hi)


def general(n, seed, i, m=None):
    from magma_amd import gen
    m = n if m is None else m
    return gen.uniform_pm1(seed, [i], n * m).reshape(n, m)


@pytest.mark.parametrize("nocheck", [False, True])
def test_getrf_vbatched(queue, oracle, nocheck):
    import magma_amd as ma
    sizes = list(sizes_cfg3(300)) + [0, 1, 2, 64, 65, 256, 300, 512]
    mats = [general(int(n), 2, i) for i, n in enumerate(sizes)]
    print(f"getrf_vbatched count={len(mats)} nocheck={nocheck}", flush=True)
    V = ma.VarBatch(mats)
    ipiv = ma.VarInts([max(int(n), 1) for n in sizes])
    info = ma.dev_ints(np.zeros(len(mats)))
    L = ma.load()
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
    for i, A in enumerate(mats):
        n = int(sizes[i])
        if n == 0:
            assert info_h[i] == 0
            continue
        LUo, ipo, info_o = oracle.dgetrf(A)
        assert info_h[i] == info_o, i
        assert np.array_equal(ipiv.numpy(i, n), ipo), i
        assert same(got[i], LUo), i


def test_getrf_vbatched_rectangular(queue, oracle):
    import magma_amd as ma
    rng = np.random.default_rng(5)
    shapes = [(int(rng.integers(1, 200)), int(rng.integers(1, 200))) for _ in range(60)]
    mats = [general(n, 9, i, m=m) for i, (m, n) in enumerate(shapes)]
    V = ma.VarBatch(mats)
    ipiv = ma.VarInts([min(m, n) for m, n in shapes])
    info = ma.dev_ints(np.zeros(len(mats)))
    ret = ma.load().magma_dgetrf_vbatched(V.d_rows.data_ptr(), V.d_cols.data_ptr(), V.ptr, V.d_ld.data_ptr(),
                                          ipiv.ptr, info.data_ptr(), V.count, queue.handle)
    assert ret == 0
    got = V.matrices()
    for i, (m, n) in enumerate(shapes):
        LUo, ipo, info_o = oracle.dgetrf(mats[i], m=m, n=n)
        assert np.array_equal(ipiv.numpy(i, min(m, n)), ipo)
        assert same(got[i], LUo)


@pytest.mark.parametrize("nocheck", [False, True])
def test_potrf_posv_vbatched(queue, oracle, nocheck):
    import magma_amd as ma
    from magma_amd import gen
    sizes = [int(n) for n in sizes_cfg3(120, seed=4, lo=1, hi=300)]
    mats = [gen.spd_batch(11, 1, n, first=i)[0] for i, n in enumerate(sizes)]
    nrhs = [1 + (i % 5) for i in range(len(sizes))]
    rhs = [gen.rhs_batch(11, 1, n, r, first=i)[0] for i, (n, r) in enumerate(zip(sizes, nrhs))]
    V = ma.VarBatch(mats)
    R = ma.VarBatch(rhs)
    info = ma.dev_ints(np.zeros(len(mats)))
    L = ma.load()
    if nocheck:
        ret = L.magma_dpotrf_vbatched_max_nocheck(ma.MagmaLower, V.d_cols.data_ptr(), V.ptr, V.d_ld.data_ptr(),
                                                  info.data_ptr(), V.count, max(sizes), queue.handle)
        queue.sync()
    else:
        ret = L.magma_dpotrf_vbatched(ma.MagmaLower, V.d_cols.data_ptr(), V.ptr, V.d_ld.data_ptr(),
                                      info.data_ptr(), V.count, queue.handle)
    assert ret == 0
    assert np.all(info.cpu().numpy()[: V.count] == 0)
    ret = L.magma_dpotrs_vbatched(ma.MagmaLower, V.d_cols.data_ptr(), R.d_cols.data_ptr(), V.ptr,
                                  V.d_ld.data_ptr(), R.ptr, R.d_ld.data_ptr(), V.count, queue.handle)
    assert ret == 0
    Lg = V.matrices()
    X = R.matrices()
    for i, n in enumerate(sizes):
        Lo, info_o = oracle.dpotrf(mats[i])
        low = np.tril(np.ones((n, n), bool)).T
        assert same(np.where(low, Lg[i], 0), np.where(low, Lo, 0)), i
        Xo = oracle.dpotrs(Lo, rhs[i])
        assert np.allclose(X[i], Xo, rtol=1e-9, atol=1e-14)
        assert oracle.solve_error(mats[i], X[i], rhs[i]) < 30


def test_getrs_vbatched(queue, oracle):
    import magma_amd as ma
    from magma_amd import gen
    sizes = [int(n) for n in sizes_cfg3(80, seed=6, lo=1, hi=200)]
    mats = [general(n, 6, i) for i, n in enumerate(sizes)]
    nrhs = [1 + (i % 3) for i in range(len(sizes))]
    rhs = [gen.rhs_batch(6, 1, n, r, first=i)[0] for i, (n, r) in enumerate(zip(sizes, nrhs))]
    V = ma.VarBatch(mats)
    R = ma.VarBatch(rhs)
    ipiv = ma.VarInts(sizes)
    info = ma.dev_ints(np.zeros(len(mats)))
    L = ma.load()
    assert L.magma_dgetrf_vbatched(V.d_rows.data_ptr(), V.d_cols.data_ptr(), V.ptr, V.d_ld.data_ptr(),
                                   ipiv.ptr, info.data_ptr(), V.count, queue.handle) == 0
    assert L.magma_dgetrs_vbatched(ma.MagmaNoTrans, V.d_cols.data_ptr(), R.d_cols.data_ptr(), V.ptr,
                                   V.d_ld.data_ptr(), ipiv.ptr, R.ptr, R.d_ld.data_ptr(), V.count,
                                   queue.handle) == 0
    X = R.matrices()
    for i, n in enumerate(sizes):
        LUo, ipo, _ = oracle.dgetrf(mats[i])
        Xo = oracle.dgetrs(LUo, ipo, rhs[i])
        assert same(X[i], Xo), i


def _gemm_ref_error(A, B, C0, Cgot, alpha, beta, oracle, m, n, k):
    Cref = oracle.dgemm(b"N", b"N", m, n, k, alpha, A, max(m, 1), B, max(k, 1), beta, C0, max(m, 1))
    An = np.linalg.norm(A) if A.size else 0.0
    Bn = np.linalg.norm(B) if B.size else 0.0
    Cn = np.linalg.norm(C0) if C0.size else 0.0
    normalize = np.sqrt(k + 2) * An * Bn + 2 * Cn
    if normalize == 0:
        normalize = 1
    return np.linalg.norm(Cgot - Cref) / normalize


@pytest.mark.parametrize("variant", ["checked", "nocheck"])
def test_gemm_vbatched(queue, oracle, variant):
    import magma_amd as ma
    from magma_amd import gen
    sizes = [int(n) for n in sizes_cfg3(200)] + [1, 7, 33, 129]
    As = [gen.uniform_pm1(2, [i], n * n).reshape(n, n) for i, n in enumerate(sizes)]
    Bs = [gen.uniform_pm1(102, [i], n * n).reshape(n, n) for i, n in enumerate(sizes)]
    Cs = [gen.uniform_pm1(202, [i], n * n).reshape(n, n) for i, n in enumerate(sizes)]
    dA, dB, dC = ma.VarBatch(As), ma.VarBatch(Bs), ma.VarBatch(Cs)
    L = ma.load()
    alpha, beta = 1.0, 1.0
    args = (ma.MagmaNoTrans, ma.MagmaNoTrans, dC.d_rows.data_ptr(), dC.d_cols.data_ptr(), dA.d_cols.data_ptr(),
            alpha, dA.ptr, dA.d_ld.data_ptr(), dB.ptr, dB.d_ld.data_ptr(), beta, dC.ptr, dC.d_ld.data_ptr(),
            dA.count)
    if variant == "checked":
        L.magmablas_dgemm_vbatched(*args, queue.handle)
    else:
        L.magmablas_dgemm_vbatched_max_nocheck(*args, max(sizes), max(sizes), max(sizes), queue.handle)
    queue.sync()
    got = dC.matrices()
    worst = 0.0
    for i, n in enumerate(sizes):
        e = _gemm_ref_error(As[i], Bs[i], Cs[i], got[i], alpha, beta, oracle, n, n, n)
        worst = max(worst, e)
    assert worst < 3 * EPS


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_gemm_batched_integer_exact(queue, oracle, ta, tb):
    import magma_amd as ma
    rng = np.random.default_rng(1)
    count, m, n, k = 40, 37, 21, 50
    sa = (k, m) if ta == "N" else (m, k)        # storage (cols, rows)
    sb = (n, k) if tb == "N" else (k, n)
    A = rng.integers(-4, 5, (count,) + sa).astype(float)
    B = rng.integers(-4, 5, (count,) + sb).astype(float)
    C = rng.integers(-4, 5, (count, n, m)).astype(float)
    dA, dB, dC = ma.DeviceBatch(A), ma.DeviceBatch(B), ma.DeviceBatch(C)
    t = {"N": ma.MagmaNoTrans, "T": ma.MagmaTrans}
    ma.load().magma_dgemm_batched(t[ta], t[tb], m, n, k, 2.0, dA.ptr, dA.ld, dB.ptr, dB.ld, -1.0,
                                  dC.ptr, dC.ld, count, queue.handle)
    queue.sync()
    got = dC.numpy()
    for s in range(count):
        ref = oracle.dgemm(ta.encode(), tb.encode(), m, n, k, 2.0, A[s], dA.ld, B[s], dB.ld, -1.0, C[s], m)
        assert np.array_equal(got[s], ref)


def test_gemm_batched_beta_zero_ignores_nan(queue):
    import magma_amd as ma
    count, n = 8, 20
    A = np.ones((count, n, n)); B = np.ones((count, n, n)); C = np.full((count, n, n), np.nan)
    dA, dB, dC = ma.DeviceBatch(A), ma.DeviceBatch(B), ma.DeviceBatch(C)
    ma.load().magmablas_dgemm_batched(ma.MagmaNoTrans, ma.MagmaNoTrans, n, n, n, 1.0, dA.ptr, n, dB.ptr, n,
                                      0.0, dC.ptr, n, count, queue.handle)
    queue.sync()
    assert np.all(dC.numpy() == n)


def test_vbatched_argument_errors(queue):
    import magma_amd as ma
    L = ma.load()
    n = ma.dev_ints([4, -1, 3])
    ld = ma.dev_ints([4, 4, 3])
    info = ma.dev_ints([0, 0, 0])
    assert L.magma_dgetrf_vbatched(n.data_ptr(), n.data_ptr(), 0, ld.data_ptr(), 0, info.data_ptr(), 3,
                                   queue.handle) == -1
    n2 = ma.dev_ints([4, 5, 3])
    assert L.magma_dgetrf_vbatched(n2.data_ptr(), n2.data_ptr(), 0, ld.data_ptr(), 0, info.data_ptr(), 3,
                                   queue.handle) == -4
    assert L.magma_dpotrf_vbatched(ma.MagmaLower, n.data_ptr(), 0, ld.data_ptr(), info.data_ptr(), 3,
                                   queue.handle) == -2
    assert L.magma_dpotrf_vbatched(ma.MagmaUpper, n2.data_ptr(), 0, ld.data_ptr(), info.data_ptr(), 3,
                                   queue.handle) == -4                # Upper is supported: ldda is checked
    assert L.magma_dpotrf_vbatched(0, n2.data_ptr(), 0, ld.data_ptr(), info.data_ptr(), 3, queue.handle) == -1


def test_gemm_tile_variants_bitwise_equal(queue, oracle):
    """Every C entry is accumulated over k in the same order whatever the tile
    (16, 32, 48 per wavefront, 64 per workgroup through LDS), so the four GEMM
    variants agree bit for bit; each also meets the tester's bound and is exact
    on integer operands in all four transpose cases."""
    import magma_amd as ma
    from magma_amd import gen
    L = ma.load()
    sizes = [int(n) for n in sizes_cfg3(120)] + [1, 7, 33, 47, 48, 49, 97, 129, 256]
    As = [gen.uniform_pm1(3, [i], n * n).reshape(n, n) for i, n in enumerate(sizes)]
    Bs = [gen.uniform_pm1(103, [i], n * n).reshape(n, n) for i, n in enumerate(sizes)]
    Cs = [gen.uniform_pm1(203, [i], n * n).reshape(n, n) for i, n in enumerate(sizes)]
    rng = np.random.default_rng(9)
    m, n, k = 61, 131, 50                                    # k: three 16-steps and a 2-wide tail
    trans = {}
    for ta in "NT":
        for tb in "NT":
            sa = (k, m) if ta == "N" else (m, k)             # storage (cols, rows)
            sb = (n, k) if tb == "N" else (k, n)
            trans[ta + tb] = (rng.integers(-4, 5, (30,) + sa).astype(float),
                              rng.integers(-4, 5, (30,) + sb).astype(float),
                              rng.integers(-4, 5, (30, n, m)).astype(float))
    t = {"N": ma.MagmaNoTrans, "T": ma.MagmaTrans}
    outs = []
    try:
        for v in range(4):
            for c in range(10):
                assert L.magma_amd_tune_set(2, c, v) == 0
            dA, dB, dC = ma.VarBatch(As), ma.VarBatch(Bs), ma.VarBatch(Cs)
            print(f"gemm variant {v}", flush=True)
            L.magmablas_dgemm_vbatched(ma.MagmaNoTrans, ma.MagmaNoTrans, dC.d_rows.data_ptr(), dC.d_cols.data_ptr(),
                                       dA.d_cols.data_ptr(), 1.0, dA.ptr, dA.d_ld.data_ptr(), dB.ptr,
                                       dB.d_ld.data_ptr(), 1.0, dC.ptr, dC.d_ld.data_ptr(), dA.count, queue.handle)
            queue.sync()
            got = dC.matrices()
            outs.append(got)
            worst = max(_gemm_ref_error(As[i], Bs[i], Cs[i], got[i], 1.0, 1.0, oracle, n, n, n)
                        for i, n in enumerate(sizes))
            assert worst < 3 * EPS, (v, worst)
            for key, (Ai, Bi, Ci) in trans.items():
                dAi, dBi, dCi = ma.DeviceBatch(Ai), ma.DeviceBatch(Bi), ma.DeviceBatch(Ci)
                L.magma_dgemm_batched(t[key[0]], t[key[1]], m, n, k, 2.0, dAi.ptr, dAi.ld, dBi.ptr,
                                      dBi.ld, -1.0, dCi.ptr, dCi.ld, 30, queue.handle)
                queue.sync()
                gi = dCi.numpy()
                for s in range(30):
                    ref = oracle.dgemm(key[0].encode(), key[1].encode(), m, n, k, 2.0, Ai[s], dAi.ld, Bi[s],
                                       dBi.ld, -1.0, Ci[s], m)
                    assert np.array_equal(gi[s], ref), (v, key, s)
    finally:
        for c in range(10):
            L.magma_amd_tune_set(2, c, -1)
    for v in (1, 2, 3):
        for i in range(len(sizes)):
            assert np.array_equal(outs[v][i], outs[0][i]), (v, sizes[i])
