"""GPU: magma_dgetrf_vbatched_max_nocheck_noalloc, the allocation-free variant
of magma_dgetrf_vbatched_max_nocheck, against that entry point.

Both calls go through the C-ABI.  The variant keeps its work list in a static
device array, so the test makes two calls in a row, the second with fewer and
different matrices, and checks factors, pivots and info bit for bit against
the allocating entry point on the same inputs each time (which is itself
pinned bitwise to the oracle by the vbatched LU tests).  A batch above the
array's limit must be rejected before anything runs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_BATCH = 1 << 20


def same(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def run(L, q, fn, sizes, seed):
    import magma_amd as ma
    from magma_amd import gen
    mats = [gen.uniform_pm1(seed, [i], n * n).reshape(n, n) for i, n in enumerate(sizes)]
    V = ma.VarBatch(mats)
    ipiv = ma.VarInts(sizes)
    info = ma.dev_ints(np.full(len(sizes), -7))
    r = getattr(L, fn)(V.d_rows.data_ptr(), V.d_cols.data_ptr(), max(sizes), V.ptr, V.d_ld.data_ptr(),
                       ipiv.ptr, info.data_ptr(), len(sizes), q.handle)
    assert r == 0, (fn, r)
    q.sync()
    return V.matrices(), ipiv.t.cpu().numpy(), info.cpu().numpy()


@pytest.mark.parametrize("seed,count,nmax", [(61, 300, 600), (62, 90, 200)])
def test_noalloc_matches_allocating_entry_point(queue, seed, count, nmax):
    import magma_amd as ma
    from magma_amd import gen
    L = ma.load()
    # two calls in a row, the second smaller: nothing of the first may leak
    for s, c, top in ((seed, count, nmax), (seed + 100, count // 3, nmax // 2)):
        sizes = [int(n) for n in gen.vbatch_sizes(s, c, 1, top)]
        print(f"noalloc getrf_vbatched seed={s} count={c} max_n={max(sizes)}", flush=True)
        A_ref, p_ref, i_ref = run(L, queue, "magma_dgetrf_vbatched_max_nocheck", sizes, s)
        A_got, p_got, i_got = run(L, queue, "magma_dgetrf_vbatched_max_nocheck_noalloc", sizes, s)
        assert np.array_equal(p_got, p_ref)
        assert np.array_equal(i_got, i_ref)
        for k in range(len(sizes)):
            assert same(A_got[k], A_ref[k]), k


def test_noalloc_rejects_batch_above_limit(queue):
    import magma_amd as ma
    L = ma.load()
    # argument 8 (batchCount) is checked before any pointer is touched
    r = L.magma_dgetrf_vbatched_max_nocheck_noalloc(1, 1, 8, 1, 1, 1, 1, MAX_BATCH + 1, queue.handle)
    assert r == -8
