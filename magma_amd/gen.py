"""Deterministic synthetic inputs for the batched factor-and-solve workloads.

Every value is a pure function of (seed, matrix index, element index):

    u = (splitmix64(seed * 2**40 + matrix * 2**20 + elem) >> 11) * 2**-53   in [0, 1)
    uniform[-1, 1] = 2u - 1
    N(0, 1) (Box-Muller) = sqrt(-2 ln(1 - u(2e))) * cos(2 pi u(2e + 1))

Matrices are column-major: element (i, j) of an n x n matrix is elem i + j*n.
Right-hand sides of a workload with matrix seed s use seed s + 100, the B and
C operands of a GEMM seeds s + 100 and s + 200.  Per-matrix scalar draws (sizes,
condition numbers) use elem = 2**20 - 1 of their matrix index.

This mirrors the role of the reference testers' lapackf77_dlarnv / magma_zmake_hpd
input generation (testing/testing_zgetrf_batched.cpp:93,133,
testing/magma_zutil.cpp:49-58) with the distributions the benchmark
configurations name.
"""
import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
SCALAR_ELEM = (1 << 20) - 1


def splitmix64(x):
    """Vectorised splitmix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def u01(seed, mat, elem):
    """u in [0,1) for broadcastable (seed, mat, elem)."""
    seed = np.asarray(seed, dtype=np.uint64)
    mat = np.asarray(mat, dtype=np.uint64)
    elem = np.asarray(elem, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = (seed << np.uint64(40)) + (mat << np.uint64(20)) + elem
    return (splitmix64(key) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def uniform_pm1(seed, mats, count):
    """(len(mats), count) array of U[-1,1] values, elem = 0..count-1."""
    mats = np.asarray(mats, dtype=np.uint64).reshape(-1, 1)
    e = np.arange(count, dtype=np.uint64).reshape(1, -1)
    return 2.0 * u01(seed, mats, e) - 1.0


def normal(seed, mats, count):
    """(len(mats), count) N(0,1) values by Box-Muller."""
    mats = np.asarray(mats, dtype=np.uint64).reshape(-1, 1)
    e = np.arange(count, dtype=np.uint64).reshape(1, -1)
    u1 = u01(seed, mats, 2 * e)
    u2 = u01(seed, mats, 2 * e + np.uint64(1))
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def scalar_u(seed, mats):
    """One u in [0,1) per matrix (sizes, condition numbers)."""
    return u01(seed, np.asarray(mats, dtype=np.uint64), np.uint64(SCALAR_ELEM))


def to_colmajor(A):
    """(b, n, n) row-major numpy matrices -> (b, n, n) arrays whose C order is column-major."""
    return np.ascontiguousarray(np.swapaxes(A, -1, -2))


def general_batch(seed, batch, n, first=0):
    """cfg1-style general matrices U[-1,1]; returns (batch, n, n) in column-major storage
    (storage[b, j, i] = A_b(i, j))."""
    mats = np.arange(first, first + batch)
    return uniform_pm1(seed, mats, n * n).reshape(batch, n, n)


def rhs_batch(seed, batch, n, nrhs, first=0):
    """U[-1,1] right-hand sides, column-major (batch, nrhs, n) storage."""
    mats = np.arange(first, first + batch)
    return uniform_pm1(seed + 100, mats, n * nrhs).reshape(batch, nrhs, n)


def spd_batch(seed, batch, n, first=0):
    """cfg2-style SPD matrices A = B B' + n I with B ~ N(0,1); column-major storage
    (symmetric, so storage order does not matter)."""
    mats = np.arange(first, first + batch)
    Bm = normal(seed, mats, n * n).reshape(batch, n, n)
    A = Bm @ np.swapaxes(Bm, -1, -2)
    A += n * np.eye(n)
    return np.ascontiguousarray(A)


def vbatch_sizes(seed, batch, lo, hi, first=0):
    """Log-uniform integer sizes in [lo, hi] (cfg3 / cfg5 rule)."""
    u = scalar_u(seed, np.arange(first, first + batch))
    n = np.floor(np.exp(np.log(lo) + u * (np.log(hi + 1) - np.log(lo)))).astype(np.int64)
    return np.clip(n, lo, hi)


def cond_batch_params(seed, batch, n, first=0):
    """cfg4: kappa = 10**(2 + 4u) and singular values s_i = kappa**(-i/(n-1))."""
    u = scalar_u(seed, np.arange(first, first + batch))
    kappa = 10.0 ** (2.0 + 4.0 * u)
    i = np.arange(n, dtype=np.float64)
    s = kappa.reshape(-1, 1) ** (-(i / (n - 1))).reshape(1, -1)
    return kappa, s


def cond_batch(seed, batch, n, first=0):
    """cfg4 matrices A = U diag(s) V' with U, V the Q factors of QR of N(0,1)
    matrices (elem 0..n^2-1 for U, n^2..2n^2-1 for V).  Returns column-major
    (batch, n, n) storage and kappa.  Uses numpy's LAPACK QR (host)."""
    mats = np.arange(first, first + batch)
    G = normal(seed, mats, 2 * n * n)
    G1 = G[:, : n * n].reshape(batch, n, n)       # storage order: column-major
    G2 = G[:, n * n:].reshape(batch, n, n)
    U, RU = np.linalg.qr(np.swapaxes(G1, -1, -2))
    V, RV = np.linalg.qr(np.swapaxes(G2, -1, -2))
    # the unique QR: R's diagonal positive
    U = U * np.sign(np.diagonal(RU, axis1=-2, axis2=-1))[..., None, :]
    V = V * np.sign(np.diagonal(RV, axis1=-2, axis2=-1))[..., None, :]
    kappa, s = cond_batch_params(seed, batch, n, first)
    A = (U * s[:, None, :]) @ np.swapaxes(V, -1, -2)
    return to_colmajor(A), kappa


# ---- the same generator on a torch device (large benchmark inputs) --------
# int64 arithmetic wraps modulo 2**64 on every torch backend; right shifts of
# negative int64 are arithmetic, so logical shifts mask the sign extension.
def _t_lsr(x, k):
    import torch
    return (x >> k) & torch.tensor((1 << (64 - k)) - 1, dtype=torch.int64, device=x.device)


def _t_u64(v):
    """python int in [0, 2**64) as the int64 with the same bits"""
    return v - (1 << 64) if v >= (1 << 63) else v


def splitmix64_torch(x):
    z = x + _t_u64(0x9E3779B97F4A7C15)
    z = (z ^ _t_lsr(z, 30)) * _t_u64(0xBF58476D1CE4E5B9)
    z = (z ^ _t_lsr(z, 27)) * _t_u64(0x94D049BB133111EB)
    return z ^ _t_lsr(z, 31)


def u01_torch(seed, mats, elems):
    """u in [0,1): mats (B,1) and elems (1,E) int64 tensors -> (B,E) float64."""
    import torch
    key = (seed << 40) + (mats << 20) + elems
    return _t_lsr(splitmix64_torch(key), 11).to(torch.float64) * 2.0 ** -53


def uniform_pm1_torch(seed, first, batch, count, device):
    import torch
    mats = torch.arange(first, first + batch, dtype=torch.int64, device=device).view(-1, 1)
    e = torch.arange(count, dtype=torch.int64, device=device).view(1, -1)
    return 2.0 * u01_torch(seed, mats, e) - 1.0


def normal_torch(seed, first, batch, count, device, elem0=0):
    import torch
    mats = torch.arange(first, first + batch, dtype=torch.int64, device=device).view(-1, 1)
    e = torch.arange(elem0, elem0 + count, dtype=torch.int64, device=device).view(1, -1)
    u1 = u01_torch(seed, mats, 2 * e)
    u2 = u01_torch(seed, mats, 2 * e + 1)
    return torch.sqrt(-2.0 * torch.log1p(-u1)) * torch.cos(2.0 * torch.pi * u2)


def _cholqr2(G):
    """Q of G = QR with R's diagonal positive, by Cholesky QR applied twice
    (Q1 = G chol(G'G)^-1, then the same on Q1): orthogonal to rounding for
    cond(G) well below 1e7, which Gaussian matrices of these sizes are."""
    import torch
    Q = G
    for _ in range(2):
        R = torch.linalg.cholesky(Q.transpose(-1, -2) @ Q, upper=True)
        Q = torch.linalg.solve_triangular(R, Q, upper=True, left=False)
    return Q


def _q_factor(G, state):
    """Q of the QR of each matrix of G (rows, cols) with R's diagonal positive,
    on G's device when the device solvers accept the batch, else on the host.
    A batched device solver can refuse a batch for want of workspace (a
    library status, not a device error); the host QR then serves the
    remaining chunks."""
    import torch
    if state.get("device_qr", True):
        try:
            return _cholqr2(G.contiguous())
        except RuntimeError:
            state["device_qr"] = False
    Q, R = torch.linalg.qr(G.cpu().contiguous())
    Q = Q * torch.sign(torch.diagonal(R, dim1=-2, dim2=-1))[..., None, :]
    return Q.to(G.device)


def cond_batch_torch(seed, batch, n, device, first=0, chunk=128, progress=None):
    """Conditioned matrices built on the device: (batch, n, n) column-major
    storage.  Same construction as cond_batch (QR of N(0,1) matrices, R's
    diagonal positive); equal to it up to rounding.  progress(done, total) is
    called after each chunk."""
    import torch
    out = torch.empty((batch, n, n), dtype=torch.float64, device=device)
    kappa_all = []
    state = {}
    for c0 in range(0, batch, chunk):
        if progress:
            progress(c0, batch)
        b = min(chunk, batch - c0)
        G1 = normal_torch(seed, first + c0, b, n * n, device).view(b, n, n)       # storage (cols, rows)
        G2 = normal_torch(seed, first + c0, b, n * n, device, elem0=n * n).view(b, n, n)
        U = _q_factor(G1.transpose(1, 2), state)
        V = _q_factor(G2.transpose(1, 2), state)
        kappa, s = cond_batch_params(seed, b, n, first + c0)
        st = torch.from_numpy(s).to(device)
        A = torch.matmul(U * st[:, None, :], V.transpose(1, 2).contiguous())     # row-major A
        out[c0:c0 + b] = A.transpose(1, 2)                                         # column-major storage
        kappa_all.append(kappa)
        del G1, G2, U, V, A
    return out, np.concatenate(kappa_all)
