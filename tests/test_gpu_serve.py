"""GPU: the request server (magma_amd_server_*): continuous batching of LU and
SPD solves under a device-memory budget with resident factorizations.

Bar: every request's solution has backward error <= n * eps64 (oracle
residual, orc_solve_error <= 1) unless it took the FP64 fallback (then < 30);
each result is bit-for-bit the same whether the request ran alone or mixed
with others arriving while it was in flight (each request runs in its own
workgroup with the same code, so the mix cannot change its arithmetic); repeat
requests with a factor key reuse the resident factorization; a budget too
small for everything evicts resident factorizations and still completes
every request.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_requests(seed, count, nmax=512, reuse_frac=0.25):
    """Request specs: (kind, n, nrhs, key, matrix index).  Keys reuse earlier
    requests' matrices with new right-hand sides."""
    from magma_amd import gen
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        if specs and rng.uniform() < reuse_frac:
            j = int(rng.integers(0, len(specs)))
            kind, n, _, key, mat = specs[j]
            if key < 0:
                key = 1000 + j
                specs[j] = (kind, n, specs[j][2], key, mat)
            specs.append((kind, n, int(rng.integers(1, 17)), key, mat))
        else:
            n = int(gen.vbatch_sizes(seed, 1, 16, nmax, first=i)[0])
            kind = int(rng.integers(0, 2))
            specs.append((kind, n, int(rng.integers(1, 17)), -1, i))
    return specs


class Fixture:
    def __init__(self, specs, seed):
        import torch
        from magma_amd import gen
        self.specs = specs
        self.A, self.B, self.X = {}, [], []
        for i, (kind, n, nrhs, key, mat) in enumerate(specs):
            if mat not in self.A:
                M = gen.spd_batch(seed, 1, n, first=mat)[0] if kind == 1 else gen.general_batch(seed, 1, n, first=mat)[0]
                self.A[mat] = torch.from_numpy(np.ascontiguousarray(M)).cuda()
            Bh = gen.rhs_batch(seed, 1, n, nrhs, first=i)[0]
            self.B.append(torch.from_numpy(np.ascontiguousarray(Bh)).cuda())
            self.X.append(torch.zeros_like(self.B[-1]))

    def request(self, i):
        from magma_amd import lib
        kind, n, nrhs, key, mat = self.specs[i]
        return lib.Request(kind, n, nrhs, self.A[mat].data_ptr(), n, self.B[i].data_ptr(), n,
                           self.X[i].data_ptr(), n, key)


def server(budget=1 << 30, max_batch=256):
    import magma_amd as ma
    L = ma.load()
    h = ctypes.c_void_p()
    assert L.magma_amd_server_create(0, budget, max_batch, ctypes.byref(h)) == 0
    return L, h


def collect(L, h, count):
    from magma_amd import lib
    buf = (lib.Completion * 4096)()
    out = {}
    while len(out) < count:
        k = L.magma_amd_server_wait(h, buf, 4096, 5.0)
        assert k >= 0
        for j in range(k):
            out[buf[j].ticket] = (buf[j].info, buf[j].iter, buf[j].reused, buf[j].berr)
    return out


def check_solutions(fx, oracle, comp, tickets):
    for i, t in enumerate(tickets):
        kind, n, nrhs, key, mat = fx.specs[i]
        info, it, reused, berr = comp[t]
        assert info == 0, (i, info)
        A = fx.A[mat].cpu().numpy()
        X = fx.X[i].cpu().numpy()
        B = fx.B[i].cpu().numpy()
        be = oracle.solve_error(A, X, B)
        assert be <= (1.0 if it >= 0 else 30.0), (i, n, kind, it, be)


def test_server_same_alone_and_in_mix(oracle):
    import torch
    specs = make_requests(41, 60)
    fx = Fixture(specs, 41)
    print(f"serve alone/mix: {len(specs)} requests", flush=True)
    L, h = server()
    try:
        # alone: one request in the server at a time
        alone_t = []
        for i in range(len(specs)):
            r = fx.request(i)
            t = L.magma_amd_server_submit(h, ctypes.byref(r))
            assert t >= 0
            assert L.magma_amd_server_drain(h) == 0
            alone_t.append(t)
        comp_alone = collect(L, h, len(specs))
        X_alone = [x.clone() for x in fx.X]
        check_solutions(fx, oracle, comp_alone, alone_t)
        for x in fx.X:
            x.zero_()
        torch.cuda.synchronize()
        # mix: everything submitted back to back, batches overlap on the device
        reqs = [fx.request(i) for i in range(len(specs))]
        mix_t = [L.magma_amd_server_submit(h, ctypes.byref(r)) for r in reqs]
        assert all(t >= 0 for t in mix_t)
        comp_mix = collect(L, h, len(specs))
        for i in range(len(specs)):
            assert torch.equal(fx.X[i], X_alone[i]), i
            assert comp_mix[mix_t[i]][1] == comp_alone[alone_t[i]][1], i          # same iteration count
        # repeat requests found their factorization resident in the alone pass
        reused = [comp_alone[alone_t[i]][2] for i in range(len(specs))]
        keyed_repeats = [i for i, s in enumerate(specs)
                         if s[3] >= 0 and any(specs[j][3] == s[3] for j in range(i))]
        assert keyed_repeats and all(reused[i] == 1 for i in keyed_repeats)
    finally:
        L.magma_amd_server_destroy(h)


def test_server_small_budget_evicts(oracle):
    from magma_amd import lib
    specs = make_requests(7, 40, nmax=300, reuse_frac=0.4)
    fx = Fixture(specs, 7)
    print(f"serve budget: {len(specs)} requests", flush=True)
    L, h = server(budget=6 << 20, max_batch=8)
    try:
        reqs = [fx.request(i) for i in range(len(specs))]
        tickets = [L.magma_amd_server_submit(h, ctypes.byref(r)) for r in reqs]
        assert all(t >= 0 for t in tickets)
        comp = collect(L, h, len(specs))
        check_solutions(fx, oracle, comp, tickets)
        st = lib.ServerStats()
        assert L.magma_amd_server_stats(h, ctypes.byref(st)) == 0
        assert st.completed == len(specs)
        assert st.bytes_peak <= st.budget_bytes
        assert st.evictions > 0
    finally:
        L.magma_amd_server_destroy(h)


def test_server_run_trace(oracle):
    from magma_amd import lib
    specs = make_requests(3, 80, nmax=256)
    fx = Fixture(specs, 3)
    rng = np.random.default_rng(3)
    arrivals = np.cumsum(rng.exponential(1.0 / 4000.0, len(specs)))
    print(f"serve trace: {len(specs)} requests over {arrivals[-1]:.3f} s", flush=True)
    L, h = server()
    try:
        reqs = (lib.Request * len(specs))(*[fx.request(i) for i in range(len(specs))])
        arr = (ctypes.c_double * len(specs))(*arrivals)
        out = (lib.Completion * len(specs))()
        assert L.magma_amd_server_run_trace(h, reqs, arr, len(specs), out) == 0
        lat = np.array([out[i].t_done - out[i].t_submit for i in range(len(specs))])
        assert np.all(lat > 0)
        comp = {out[i].ticket: (out[i].info, out[i].iter, out[i].reused, out[i].berr) for i in range(len(specs))}
        check_solutions(fx, oracle, comp, [out[i].ticket for i in range(len(specs))])
    finally:
        L.magma_amd_server_destroy(h)


def test_server_rejects_bad_requests():
    from magma_amd import lib
    L, h = server(budget=1 << 20)
    try:
        r = lib.Request(0, 600, 1, 1, 600, 1, 600, 1, 600, -1)       # n above 512
        assert L.magma_amd_server_submit(h, ctypes.byref(r)) == -2
        r = lib.Request(5, 10, 1, 1, 10, 1, 10, 1, 10, -1)           # unknown kind
        assert L.magma_amd_server_submit(h, ctypes.byref(r)) == -1
        r = lib.Request(0, 10, 1, 1, 9, 1, 10, 1, 10, -1)            # ldda < n
        assert L.magma_amd_server_submit(h, ctypes.byref(r)) == -4
        r = lib.Request(0, 400, 1, 1, 400, 1, 400, 1, 400, -1)       # does not fit a 1 MiB budget
        assert L.magma_amd_server_submit(h, ctypes.byref(r)) == lib.MAGMA_ERR_DEVICE_ALLOC
    finally:
        L.magma_amd_server_destroy(h)
