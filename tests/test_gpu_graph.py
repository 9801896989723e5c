"""GPU: vbatched factor + solve captured in a HIP graph and replayed.

The *_max_nocheck vbatched entry points never synchronize, never allocate
once the queue's workspace has been sized by an eager call of the same batch
count, clear their device work lists with kernels (no memset nodes) and read
their size-class counts on the device, so getrf/potrf -> getrs/potrs can be
captured once and replayed.  Between replays the test runs unrelated eager
work and restores the inputs with ordinary copies; every replay must
reproduce the eager result bit for bit, and the eager result must match the
oracle bit for bit (factors, pivots, info, no-transpose / forward solves).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def same(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


@pytest.mark.parametrize("kind", ["lu", "chol"])
def test_vbatched_factor_solve_graph_replay(oracle, kind):
    import torch
    import magma_amd as ma
    from magma_amd import gen

    count = 160
    sizes = [int(n) for n in gen.vbatch_sizes(21, count, 4, 300)]
    nrhs = [1 + (i % 4) for i in range(count)]
    if kind == "lu":
        mats = [gen.uniform_pm1(21, [i], n * n).reshape(n, n) for i, n in enumerate(sizes)]
    else:
        mats = [gen.spd_batch(21, 1, n, first=i)[0] for i, n in enumerate(sizes)]
    rhs = [gen.rhs_batch(21, 1, n, r, first=i)[0] for i, (n, r) in enumerate(zip(sizes, nrhs))]
    print(f"graph replay kind={kind} count={count} max_n={max(sizes)}", flush=True)

    s = torch.cuda.Stream()
    q = ma.Queue(0, stream=s)
    L = ma.load()
    L.magma_queue_set_async(q.handle, 1)
    with torch.cuda.stream(s):
        V = ma.VarBatch(mats)
        R = ma.VarBatch(rhs)
        ipiv = ma.VarInts(sizes)
        info = ma.dev_ints(np.zeros(count))
        A0 = V.t.clone()
        B0 = R.t.clone()
    max_n = max(sizes)

    def calls():
        if kind == "lu":
            r1 = L.magma_dgetrf_vbatched_max_nocheck(V.d_rows.data_ptr(), V.d_cols.data_ptr(), max_n, V.ptr,
                                                     V.d_ld.data_ptr(), ipiv.ptr, info.data_ptr(), count,
                                                     q.handle)
            r2 = L.magma_dgetrs_vbatched_max_nocheck(ma.MagmaNoTrans, V.d_cols.data_ptr(), R.d_cols.data_ptr(),
                                                     V.ptr, V.d_ld.data_ptr(), ipiv.ptr, R.ptr,
                                                     R.d_ld.data_ptr(), count, max_n, q.handle)
        else:
            r1 = L.magma_dpotrf_vbatched_max_nocheck(ma.MagmaLower, V.d_cols.data_ptr(), V.ptr,
                                                     V.d_ld.data_ptr(), info.data_ptr(), count, max_n,
                                                     q.handle)
            r2 = L.magma_dpotrs_vbatched_max_nocheck(ma.MagmaLower, V.d_cols.data_ptr(), R.d_cols.data_ptr(),
                                                     V.ptr, V.d_ld.data_ptr(), R.ptr, R.d_ld.data_ptr(), count,
                                                     max_n, q.handle)
        assert r1 == 0 and r2 == 0

    # eager run (also sizes the queue workspace)
    with torch.cuda.stream(s):
        calls()
    s.synchronize()
    LU_e, X_e = V.t.clone(), R.t.clone()
    ipiv_e, info_e = ipiv.t.clone(), info.clone()

    # the eager result is the oracle's
    got_A, got_X = V.matrices(), R.matrices()
    info_h = info.cpu().numpy()
    for i, n in enumerate(sizes):
        if kind == "lu":
            Fo, ipo, info_o = oracle.dgetrf(mats[i])
            Xo = oracle.dgetrs(Fo, ipo, rhs[i])
            assert np.array_equal(ipiv.numpy(i, n), ipo), i
            assert same(got_A[i], Fo), i
        else:
            Fo, info_o = oracle.dpotrf(mats[i])
            Xo = oracle.dpotrs(Fo, rhs[i])
            assert same(np.tril(got_A[i].T), np.tril(Fo.T)), i
        assert info_h[i] == info_o, i
        if kind == "lu":
            assert same(got_X[i], Xo), i
        else:
            assert oracle.solve_error(mats[i], got_X[i], rhs[i]) < 30, i

    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        V.t.copy_(A0)
        R.t.copy_(B0)
    s.synchronize()
    with torch.cuda.graph(g, stream=s):
        calls()
    for rep in range(3):
        with torch.cuda.stream(s):
            V.t.copy_(A0)
            R.t.copy_(B0)
            ipiv.t.fill_(-7)
            info.fill_(-7)
            # unrelated eager work between replays
            z = torch.randn(300, 300, dtype=torch.float64, device="cuda")
            _ = (z @ z).sum().item()
            g.replay()
        s.synchronize()
        assert torch.equal(V.t, LU_e), rep
        assert torch.equal(R.t, X_e), rep
        if kind == "lu":
            assert torch.equal(ipiv.t, ipiv_e), rep
        assert torch.equal(info[:count], info_e[:count]), rep       # the extra slot is library scratch
    q.close()
