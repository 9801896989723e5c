"""GPU parity: magma_dpotrf/dpotrs/dposv_batched against the CPU oracle.

Bar: info bit-exact; the Cholesky factor bit-exact (same update order and
fused multiply-adds as the oracle); the upper triangle of A untouched; the
forward solve bit-exact and the backward (L') solve to rounding, checked by
the tester residual < 30 and the factor residual ||A - LL'||/(n||A||eps) < 30.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 7, 8, 15, 16, 17, 31, 32, 33, 47, 64, 65, 100, 128, 129, 256, 300]


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
    count = 200 if n <= 64 else 30
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


@pytest.mark.parametrize("n", [5, 40, 64, 150])
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


def test_potrf_upper_rejected(queue):
    import magma_amd as ma
    L = ma.load()
    assert L.magma_dpotrf_batched(ma.MagmaUpper, 4, 0, 4, 0, 1, queue.handle) == -1
    assert L.magma_dpotrf_batched(ma.MagmaLower, -1, 0, 4, 0, 1, queue.handle) == -2
    assert L.magma_dpotrf_batched(ma.MagmaLower, 4, 0, 3, 0, 1, queue.handle) == -4
