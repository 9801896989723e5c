"""GPU: every kernel variant the autotuner can select, each forced through
magma_amd_tune_set and checked to the bar of its entry point, and the
autotuned table itself on the benchmark's conditioned n = 256 inputs.

  LU (family 0, classes 4-6: orders up to 128, 256, 512): pivots, info and
  factors bit-exact against the oracle's netlib-order dgetrf, on uniform
  matrices and on the conditioned U diag(s) V' matrices the mixed benchmark
  uses.  The register-resident variants of class 4 (lu_reg128.hip) also on
  rectangular and padded shapes, through the vbatched entry point, and on
  integer matrices with zero columns, exact ties, NaN and Inf entries (the
  last two take its exact path part-way through the factorization).
  Mixed solver (family 1, classes 4-6): dsgesv and dsposv backward error
  ||b - Ax||_inf / (||A||_inf ||x||_inf) <= n eps64 for every converged matrix
  (orc_solve_error <= 1), the stopping test of src/zcgesv_gpu.cpp:266-279;
  matrices that fall back (singular, cond 1e13) hold the FP64 factors and
  solutions bit-exact as the plain path produces them.

The variant count of each family is probed through magma_amd_tune_set, which
rejects a variant number past the last, so a new variant is covered here as
soon as the tuner can pick it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GETRF, MIXED = 0, 1


def _variants(family):
    import magma_amd as ma
    L = ma.load()
    out = []
    for v in range(16):
        if L.magma_amd_tune_set(family, 5, v) != 0:
            break
        out.append(v)
    L.magma_amd_tune_set(family, 5, -1)
    return out


@pytest.fixture
def table():
    import magma_amd as ma
    L = ma.load()
    yield L
    for fam in range(3):
        for c in range(10):
            L.magma_amd_tune_set(fam, c, -1)


def _same(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


LU_CASES = [(c, v, n) for c, ns in ((4, (65, 100, 128)), (5, (129, 200, 256)), (6, (257, 300, 512)))
            for v in range(5 if c == 4 else 3) for n in ns]


@pytest.mark.parametrize("cls,variant,n", LU_CASES)
def test_getrf_variant_bitexact(queue, oracle, table, cls, variant, n):
    import magma_amd as ma
    from magma_amd import gen
    assert table.magma_amd_tune_set(GETRF, cls, variant) == 0
    count = 12
    A = gen.uniform_pm1(500 + n, np.arange(count), n * n).reshape(count, n, n).copy()
    A[3][:, 7 % n] = 0.0                                  # zero column: info, factorization continues
    A[5][(11 * n) // 13, :] = np.nan if n % 2 else 1.0    # NaN row (odd n) / ties (even n)
    print(f"getrf variant class={cls} v={variant} n={n}", flush=True)
    dA = ma.DeviceBatch(A)
    ret, ipiv, info = ma.dgetrf_batched(dA, queue)
    assert ret == 0
    LUo, ipo, info_o = oracle.batch_dgetrf(A)
    assert np.array_equal(info.cpu().numpy(), info_o)
    assert np.array_equal(ipiv.numpy(), ipo)
    assert _same(dA.numpy(), LUo)


REG128_SHAPES = [(65, 65, None), (128, 128, None), (128, 100, None), (100, 128, None), (128, 9, None),
                 (9, 128, None), (70, 127, 75), (127, 70, 130), (128, 128, 131)]


@pytest.mark.parametrize("variant", [3, 4])
@pytest.mark.parametrize("m,n,ld", REG128_SHAPES)
def test_getrf_reg128_shapes(queue, oracle, table, variant, m, n, ld):
    import magma_amd as ma
    from magma_amd import gen
    assert table.magma_amd_tune_set(GETRF, 4, variant) == 0
    count = 20
    A = gen.uniform_pm1(900 + m + n, np.arange(count), m * n).reshape(count, n, m).copy()
    print(f"getrf reg128 v={variant} m={m} n={n} ld={ld}", flush=True)
    dA = ma.DeviceBatch(A, ld=ld)
    ret, ipiv, info = ma.dgetrf_batched(dA, queue, m=m, n=n)
    assert ret == 0
    got, ip, inf = dA.numpy(), ipiv.numpy(), info.cpu().numpy()
    for s in range(count):
        LUo, ipo, info_o = oracle.dgetrf(A[s], m=m, n=n)
        assert inf[s] == info_o, s
        assert np.array_equal(ip[s], ipo), s
        assert _same(got[s], LUo), s


@pytest.mark.parametrize("variant", [3, 4])
def test_getrf_reg128_special(queue, oracle, table, variant):
    """Integer entries (exact ties in the pivot search), zero columns (info,
    the factorization continues), duplicated rows, NaN and Inf entries at
    different depths, an all-zero and a permutation matrix."""
    import magma_amd as ma
    assert table.magma_amd_tune_set(GETRF, 4, variant) == 0
    n = 97
    rng = np.random.default_rng(97)
    mats = []
    for t in range(30):
        A = rng.integers(-2, 3, size=(n, n)).astype(float)
        if t % 6 == 1:
            A[:, (5 * t) % n] = 0.0
        if t % 6 == 2:
            A[(t + 1) % n, :] = A[t % n, :]
        if t % 6 == 3:
            A[(t * 7) % n, (t * 3) % n] = np.nan
        if t % 6 == 4:
            A[(t * 11) % n, (t * 5) % n] = np.inf
            A[(t * 13) % n, (t * 5 + 40) % n] = -np.inf
        if t % 6 == 5:
            A = np.abs(A) * 1e300                     # products overflow to Inf during the update
        mats.append(A.T.copy())
    mats.append(np.zeros((n, n)))
    mats.append(np.eye(n)[::-1].T.copy())
    A = np.stack(mats)
    print(f"getrf reg128 special v={variant}", flush=True)
    dA = ma.DeviceBatch(A)
    ret, ipiv, info = ma.dgetrf_batched(dA, queue)
    assert ret == 0
    LUo, ipo, info_o = oracle.batch_dgetrf(A)
    assert np.array_equal(info.cpu().numpy(), info_o)
    assert np.array_equal(ipiv.numpy(), ipo)
    assert _same(dA.numpy(), LUo)


@pytest.mark.parametrize("variant", [3, 4])
def test_getrf_reg128_vbatched(queue, oracle, table, variant):
    import magma_amd as ma
    from magma_amd import gen
    assert table.magma_amd_tune_set(GETRF, 4, variant) == 0
    rng = np.random.default_rng(128 + variant)
    shapes = [(int(rng.integers(1, 129)), int(rng.integers(65, 129))) for _ in range(40)]
    shapes = [(m, n) if i % 2 else (n, m) for i, (m, n) in enumerate(shapes)]
    mats = [gen.uniform_pm1(77, [i], m * n).reshape(n, m).copy() for i, (m, n) in enumerate(shapes)]
    print(f"getrf reg128 vbatched v={variant}", flush=True)
    V = ma.VarBatch(mats)
    ipiv = ma.VarInts([min(m, n) for m, n in shapes])
    info = ma.dev_ints(np.zeros(len(mats)))
    ret = table.magma_dgetrf_vbatched(V.d_rows.data_ptr(), V.d_cols.data_ptr(), V.ptr, V.d_ld.data_ptr(),
                                      ipiv.ptr, info.data_ptr(), V.count, queue.handle)
    assert ret == 0
    got = V.matrices()
    inf = info.cpu().numpy()
    for i, (m, n) in enumerate(shapes):
        LUo, ipo, info_o = oracle.dgetrf(mats[i], m=m, n=n)
        assert inf[i] == info_o, i
        assert np.array_equal(ipiv.numpy(i, min(m, n)), ipo), i
        assert _same(got[i], LUo), i


@pytest.mark.parametrize("variant", range(3))
def test_getrf_variant_cond256(queue, oracle, table, variant):
    """The benchmark's plain path on its own inputs: 16 conditioned matrices
    of order 256 (seed 3), pivots and factors bit-exact, and the solve."""
    import magma_amd as ma
    from magma_amd import gen
    assert table.magma_amd_tune_set(GETRF, 5, variant) == 0
    n, count = 256, 16
    A, _ = gen.cond_batch(3, count, n)
    B = gen.rhs_batch(3, count, n, 1)
    print(f"getrf+getrs cond256 v={variant}", flush=True)
    dA, dB = ma.DeviceBatch(A), ma.DeviceBatch(B)
    ret, ipiv, info = ma.dgetrf_batched(dA, queue)
    assert ret == 0
    assert ma.dgetrs_batched(dA, ipiv, dB, queue) == 0
    LUo, ipo, Xo, info_o = oracle.batch_dgesv(A, B)
    assert np.array_equal(info.cpu().numpy(), info_o)
    assert np.array_equal(ipiv.numpy(), ipo)
    assert _same(dA.numpy(), LUo)
    assert _same(dB.numpy(), Xo)


def _run_mixed(queue, A, B, kind):
    import magma_amd as ma
    import torch
    count, n, _ = A.shape
    nrhs = B.shape[1]
    dA, dB, dX = ma.DeviceBatch(A), ma.DeviceBatch(B), ma.DeviceBatch(np.zeros_like(B))
    ipiv = ma.IntBatch(count, n)
    info = torch.zeros(count, dtype=torch.int32, device="cuda")
    it = torch.full((count,), 99, dtype=torch.int32, device="cuda")
    L = ma.load()
    if kind == "L":
        ret = L.magma_dsgesv_batched(n, nrhs, dA.ptr, dA.ld, ipiv.ptr, dB.ptr, dB.ld, dX.ptr, dX.ld,
                                     info.data_ptr(), it.data_ptr(), count, queue.handle)
    else:
        ret = L.magma_dsposv_batched(ma.MagmaLower, n, nrhs, dA.ptr, dA.ld, dB.ptr, dB.ld, dX.ptr, dX.ld,
                                     info.data_ptr(), it.data_ptr(), count, queue.handle)
    assert ret == 0
    return dA.numpy(), dX.numpy(), ipiv.numpy(), info.cpu().numpy(), it.cpu().numpy()


def _cond13(n, seed):
    U, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((n, n)))
    V, _ = np.linalg.qr(np.random.default_rng(seed + 1).standard_normal((n, n)))
    return np.ascontiguousarray(((U * 10.0 ** (-13 * np.arange(n) / (n - 1))) @ V.T).T)


def _check_dsgesv(oracle, A, B, Ag, X, ipiv, info, it, fallback):
    n = A.shape[1]
    for s in range(A.shape[0]):
        if s in fallback:
            assert it[s] < 0, (s, it[s])
            LUo, ipo, info_o = oracle.dgetrf(A[s])
            assert info[s] == info_o
            assert np.array_equal(ipiv[s], ipo), s
            assert _same(Ag[s], LUo), s
            if info_o == 0:
                assert _same(X[s], oracle.dgetrs(LUo, ipo, B[s])), s
            continue
        assert info[s] == 0 and 0 <= it[s] <= 10, (s, it[s])
        assert np.array_equal(Ag[s], A[s])                      # untouched when refinement succeeds
        be = oracle.solve_error(A[s], X[s], B[s])               # ||r|| / (n ||A|| ||x|| eps)
        assert be <= 1.0, (s, be, it[s])


MIXED_N = {4: (70, 128), 5: (129, 256), 6: (300, 512)}
# variants of the mixed family per size class: five kernels for orders up to
# 256 (classes 4 and 5), two for orders up to 512
MIXED_VARIANTS = {4: 5, 5: 5, 6: 2}


def _mixed_cases():
    cases = []
    for cls, ns in MIXED_N.items():
        for v in range(MIXED_VARIANTS[cls]):
            for n in ns:
                cases.append((cls, v, n))
    return cases


@pytest.mark.parametrize("cls,variant,n", _mixed_cases())
def test_dsgesv_variant(queue, oracle, table, cls, variant, n):
    from magma_amd import gen
    assert table.magma_amd_tune_set(MIXED, cls, variant) == 0
    count = 20 if n <= 256 else 8
    nrhs = 2 if n % 2 else 1
    A, _ = gen.cond_batch(3, count, n)
    A = A.copy()
    A[1] = 0.0                                     # singular: iter -3, FP64 path with info 1
    A[2] = _cond13(n, 7)                           # beyond refinement: iter -11, FP64 path
    B = gen.rhs_batch(3, count, n, nrhs)
    print(f"dsgesv variant class={cls} v={variant} n={n}", flush=True)
    Ag, X, ipiv, info, it = _run_mixed(queue, A, B, "L")
    assert it[1] == -3 and info[1] == 1
    assert it[2] == -11
    _check_dsgesv(oracle, A, B, Ag, X, ipiv, info, it, fallback={1, 2})


@pytest.mark.parametrize("cls,variant,n", _mixed_cases())
def test_dsposv_variant(queue, oracle, table, cls, variant, n):
    from magma_amd import gen
    assert table.magma_amd_tune_set(MIXED, cls, variant) == 0
    count = 12 if n <= 256 else 6
    nrhs = 3 if n % 2 else 1
    A = gen.spd_batch(4, count, n).copy()
    A[1][n // 3, n // 3] = -1.0                    # not SPD: FP64 potrf reports info
    B = gen.rhs_batch(4, count, n, nrhs)
    print(f"dsposv variant class={cls} v={variant} n={n}", flush=True)
    Ag, X, _, info, it = _run_mixed(queue, A, B, "C")
    _, info_o = oracle.dpotrf(A[1])
    assert it[1] < 0 and info[1] == info_o > 0
    for s in range(count):
        if s == 1:
            continue
        assert info[s] == 0 and 0 <= it[s] <= 10, (s, it[s])
        assert np.array_equal(Ag[s], A[s])
        assert oracle.solve_error(A[s], X[s], B[s]) <= 1.0, s


def test_autotuned_dsgesv_on_cond256(queue, oracle, table):
    """magma_amd_autotune fills the table on the device; the tuned dsgesv then
    solves 64 of the benchmark's conditioned matrices (seed 3, n = 256, cond
    log-uniform in [1e2, 1e6]) to backward error <= n eps each, and the
    tuned plain LU gives the oracle's pivots and factors bit for bit."""
    import magma_amd as ma
    from magma_amd import gen
    print("autotune", flush=True)
    assert table.magma_amd_autotune(queue.handle, 0) == 0
    chosen = [table.magma_amd_tune_get(f, c) for f in (GETRF, MIXED) for c in (4, 5, 6)]
    print("tuned getrf/mixed classes 4-6:", chosen, flush=True)
    assert all(v >= 0 for v in chosen)
    n, count = 256, 64
    A, _ = gen.cond_batch(3, count, n)
    B = gen.rhs_batch(3, count, n, 1)
    Ag, X, ipiv, info, it = _run_mixed(queue, A, B, "L")
    _check_dsgesv(oracle, A, B, Ag, X, ipiv, info, it, fallback=set(np.nonzero(it < 0)[0]))
    assert np.sum(it < 0) == 0
    dA = ma.DeviceBatch(A[:16])
    ret, ipv, inf = ma.dgetrf_batched(dA, queue)
    LUo, ipo, info_o = oracle.batch_dgetrf(A[:16])
    assert ret == 0 and np.array_equal(ipv.numpy(), ipo) and _same(dA.numpy(), LUo)
