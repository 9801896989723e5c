#!/usr/bin/env python3
"""Benchmark of the batched small-matrix factor-and-solve path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints ONE JSON line.  A step is one call sequence of the library's C API over
one batch of synthetic inputs that are already resident in HBM (each step gets
its own pre-staged copy, since factorizations work in place).

Workloads (named by what they are):
  cond256_mixed_vs_plain  8192 matrices n=256, cond log-uniform in [1e2, 1e6],
                          nrhs=1: plain FP64 (magma_dgetrf_batched +
                          magma_dgetrs_batched) and mixed precision
                          (magma_dsgesv_batched); value = mixed solves/s
  lu32_gesv               10000 general n=32, nrhs=1, magma_dgesv_batched
  chol64_posv             20000 SPD n=64, nrhs=4, magma_dposv_batched
  vbatch_lu_gemm          50000 log-uniform n in [8,256]: magma_dgetrf_vbatched
                          + magmablas_dgemm_vbatched (C += A B)

The roofline object covers the dominant kernel (the factorization), timed with
HIP events on the queue's stream; the cpu_baseline object times the CPU oracle
(the reference's CPU path restated: LAPACK-semantics routines in C looped over
the batch under OpenMP) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s at FP64 backward err ≤ n·eps (plain vs mixed-prec) + p50/p99 latency"
HBM_PEAK = 8.0e12           # B/s
FP64_PEAK = 78.6e12         # flop/s (vector and matrix FP64 on MI355X)
EPS = 2.0 ** -53


def getrf_flops(n):
    return 2.0 * n ** 3 / 3.0 - n * n / 2.0 + 5.0 * n / 6.0


def getrf_bytes(n):
    return 16.0 * n * n + 4.0 * n + 4.0


def getrs_bytes(n, r):
    return 8.0 * n * n + 4.0 * n + 16.0 * n * r


def potrf_flops(n):
    return n ** 3 / 3.0 + n * n / 2.0 + n / 6.0


def potrf_bytes(n):
    return 8.0 * n * (n + 1) + 4.0


def potrs_bytes(n, r):
    return 4.0 * n * (n + 1) + 16.0 * n * r


class Timer:
    """HIP events recorded on the library queue's stream (torch's current stream)."""

    def __init__(self, torch):
        self.torch = torch
        self.pairs = []

    def start(self):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record()
        return e

    def stop(self, e0):
        e1 = self.torch.cuda.Event(enable_timing=True)
        e1.record()
        self.pairs.append((e0, e1))

    def ms(self):
        return [a.elapsed_time(b) for a, b in self.pairs]


def measured_traffic(workload, kernel, algorithmic_bytes):
    """HBM bytes per launch of `kernel` measured by rocprofv3 --pmc FETCH_SIZE
    and WRITE_SIZE (separate runs), each scaled by the same counter's reading
    on a calibration kernel of known traffic; read from the committed profile
    summary (profiles/hbm_traffic.json).  {} when none is recorded."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic.json")
    try:
        with open(path) as fh:
            rec = json.load(fh)[workload][kernel]
    except (OSError, KeyError, ValueError):
        return {}
    t = rec["fetch_kib"] * 1024 * rec["fetch_scale"] + rec["write_kib"] * 1024 * rec["write_scale"]
    return {"traffic": t, "traffic_over_algorithmic": t / algorithmic_bytes,
            "traffic_source": "profiles/hbm_traffic.json (" + rec["source"] + ")"}


def log(msg):
    """Progress line on stderr (stdout carries only the JSON result)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def cpu_threads():
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


# ---------------------------------------------------------------------------
def run_fixed(args, torch, ma, L, q, make_inputs, step_fn, solves_per_step, kernel_units, unit_label):
    """Generic driver: stage warmup+steps copies, time steps, return stats."""
    copies = [make_inputs(i) for i in range(args.warmup + args.steps)]
    torch.cuda.synchronize()
    log(f"staged {len(copies)} input copies; warmup {args.warmup}, timed {args.steps}")
    for i in range(args.warmup):
        step_fn(copies[i], None)
    torch.cuda.synchronize()
    timer = Timer(torch)
    t0 = time.perf_counter()
    for i in range(args.steps):
        step_fn(copies[args.warmup + i], timer)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = timer.ms()
    log(f"timed region {wall * 1e3:.2f} ms")
    return wall, step_ms, copies


def workload_lu32(args, torch, ma, L, q):
    from magma_amd import gen
    count, n, nrhs = 10000, 32, 1
    A0 = torch.from_numpy(gen.general_batch(0, count, n)).cuda()
    B0 = torch.from_numpy(gen.rhs_batch(0, count, n, nrhs)).cuda()
    ipiv = ma.IntBatch(count, n)
    info = torch.zeros(count, dtype=torch.int32, device="cuda")

    def make(i):
        A = A0.clone(); B = B0.clone()
        return (A, B, ma.ptr_array(A, count, n * n), ma.ptr_array(B, count, n * nrhs))

    def step(c, timer):
        A, B, pa, pb = c
        e = timer.start() if timer else None
        L.magma_dgesv_batched(n, nrhs, pa.data_ptr(), n, ipiv.ptr, pb.data_ptr(), n, info.data_ptr(), count,
                              q.handle)
        if timer:
            timer.stop(e)

    wall, step_ms, copies = run_fixed(args, torch, ma, L, q, make, step, count, count, "matrices")
    # parity spot check of the last step against the oracle on a sample
    check = {}
    if args.check:
        from oracle import pyoracle
        A_h = A0[:200].cpu().numpy(); B_h = B0[:200].cpu().numpy()
        LUo, ipo, Xo, info_o = pyoracle.batch_dgesv(A_h, B_h)
        X = copies[-1][1][:200].cpu().numpy()
        check = {"sample": 200, "ipiv_equal": bool(np.array_equal(ipiv.numpy()[:200], ipo)),
                 "x_bitexact": bool(np.array_equal(X, Xo)),
                 "max_backward_err_over_n_eps": float(max(pyoracle.solve_error(A_h[s], X[s], B_h[s])
                                                          for s in range(0, 200, 20)))}
    ms = wall / args.steps * 1e3
    kern_ms = float(np.mean(step_ms))
    byt = count * (getrf_bytes(n) + 16.0 * n * nrhs)     # fused factor + solve
    return dict(value=count / (ms * 1e-3), unit="solves/s", ms=ms, dtype="f64",
                config={"workload": "lu32_gesv", "batch": count, "n": n, "nrhs": nrhs,
                        "api": "magma_dgesv_batched", "inputs": "U[-1,1] seed 0"},
                roofline={"bound": "hbm", "achieved": byt / (kern_ms * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
                          "unit": "GB/s", "kernel": "getrf_reg_kernel<32, fused solve>",
                          "kernel_ms": kern_ms, "bytes_per_launch": byt},
                check=check, step_ms=step_ms,
                cpu=dict(kind="dgesv", n=n, nrhs=nrhs, gen=lambda k: (gen.general_batch(0, k, n),
                                                                     gen.rhs_batch(0, k, n, nrhs))))


def workload_chol64(args, torch, ma, L, q):
    from magma_amd import gen
    count, n, nrhs = 20000, 64, 4
    A0 = torch.from_numpy(gen.spd_batch(1, count, n)).cuda()
    B0 = torch.from_numpy(gen.rhs_batch(1, count, n, nrhs)).cuda()
    info = torch.zeros(count, dtype=torch.int32, device="cuda")

    def make(i):
        A = A0.clone(); B = B0.clone()
        return (A, B, ma.ptr_array(A, count, n * n), ma.ptr_array(B, count, n * nrhs))

    def step(c, timer):
        A, B, pa, pb = c
        e = timer.start() if timer else None
        L.magma_dposv_batched(ma.MagmaLower, n, nrhs, pa.data_ptr(), n, pb.data_ptr(), n, info.data_ptr(),
                              count, q.handle)
        if timer:
            timer.stop(e)

    wall, step_ms, copies = run_fixed(args, torch, ma, L, q, make, step, count, count, "matrices")
    check = {}
    if args.check:
        from oracle import pyoracle
        A_h = A0[:100].cpu().numpy(); B_h = B0[:100].cpu().numpy()
        X = copies[-1][1][:100].cpu().numpy()
        check = {"sample": 100, "info_all_zero": bool(int(info.abs().sum()) == 0),
                 "max_backward_err_over_n_eps": float(max(pyoracle.solve_error(A_h[s], X[s], B_h[s])
                                                          for s in range(0, 100, 10)))}
    ms = wall / args.steps * 1e3
    kern_ms = float(np.mean(step_ms))
    byt = count * (potrf_bytes(n) + 16.0 * n * nrhs)
    return dict(value=count / (ms * 1e-3), unit="solves/s", ms=ms, dtype="f64",
                config={"workload": "chol64_posv", "batch": count, "n": n, "nrhs": nrhs,
                        "api": "magma_dposv_batched", "inputs": "B B'+64 I, B~N(0,1) seed 1"},
                roofline={"bound": "hbm", "achieved": byt / (kern_ms * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
                          "unit": "GB/s", "kernel": "potrf_reg_kernel<64, fused solve>",
                          "kernel_ms": kern_ms, "bytes_per_launch": byt},
                check=check, step_ms=step_ms,
                cpu=dict(kind="dposv", n=n, nrhs=nrhs, gen=lambda k: (gen.spd_batch(1, k, n),
                                                                     gen.rhs_batch(1, k, n, nrhs))))


def gpu_backward_errors(torch, A, X, B, chunk=512):
    """Per-matrix FP64 backward error ||b - A x||_inf / (||A||_inf ||x||_inf)
    divided by n eps64, computed on the device for every matrix and every
    right-hand side (A: (count, n, n) column-major storage, [s, j, i] = A(i, j);
    X, B: (count, nrhs, n)).  Returns a float64 numpy array (count,)."""
    count, n, _ = A.shape
    out = []
    for c0 in range(0, count, chunk):
        a = A[c0:c0 + chunk]
        x = X[c0:c0 + chunk]
        b = B[c0:c0 + chunk]
        # (x' A')' = A x: row vectors times the column-major storage
        ax = torch.bmm(x.contiguous(), a.contiguous())
        r = (b - ax).abs().amax(dim=2)
        anrm = a.abs().sum(dim=1).amax(dim=1, keepdim=True)        # max row sum
        xn = x.abs().amax(dim=2)
        be = r / (anrm * xn) / (n * EPS)
        out.append(be.amax(dim=1))
    return torch.cat(out).cpu().numpy()


def workload_cond256(args, torch, ma, L, q):
    """8192 matrices of order 256 with cond log-uniform in [1e2, 1e6], one
    right-hand side: plain FP64 getrf + getrs against mixed dsgesv."""
    from magma_amd import gen
    count, n, nrhs = args.batch or 8192, 256, 1
    log(f"generating {count} conditioned matrices n={n}")
    A0, kappa = gen.cond_batch_torch(
        3, count, n, "cuda", progress=lambda d, t: d % 2048 == 0 and log(f"  generated {d}/{t}"))
    log("inputs ready")
    B0 = gen.uniform_pm1_torch(103, 0, count, n * nrhs, "cuda").view(count, nrhs, n).contiguous()
    ipiv = ma.IntBatch(count, n)
    info = torch.zeros(count, dtype=torch.int32, device="cuda")
    fact = Timer(torch)
    step_t = Timer(torch)

    def make(i):
        A = A0.clone(); B = B0.clone()
        return (A, B, ma.ptr_array(A, count, n * n), ma.ptr_array(B, count, n * nrhs))

    def step_plain(c, timer):
        A, B, pa, pb = c
        e0 = step_t.start() if timer else None
        e = fact.start() if timer else None
        L.magma_dgetrf_batched(n, n, pa.data_ptr(), n, ipiv.ptr, info.data_ptr(), count, q.handle)
        if timer:
            fact.stop(e)
        L.magma_dgetrs_batched(ma.MagmaNoTrans, n, nrhs, pa.data_ptr(), n, ipiv.ptr, pb.data_ptr(), n, count,
                               q.handle)
        if timer:
            step_t.stop(e0)

    res = {}
    wall, _, copies = run_fixed(args, torch, ma, L, q, make, step_plain, count, count, "matrices")
    plain_ms = wall / args.steps * 1e3
    fact_ms = float(np.mean(fact.ms()))
    plain_steps = np.asarray(step_t.ms())
    res["plain"] = dict(solves_per_s=count / (plain_ms * 1e-3), ms_per_step=plain_ms, factor_kernel_ms=fact_ms,
                        step_p50_ms=float(np.percentile(plain_steps, 50)),
                        step_p99_ms=float(np.percentile(plain_steps, 99)))
    check = {}
    # every plain solution: FP64 backward error on the device
    Xp = copies[-1][1]
    be_p = gpu_backward_errors(torch, A0, Xp, B0)
    check["plain_backward_err_over_n_eps_max"] = float(be_p.max())
    check["plain_backward_err_above_n_eps"] = int(np.sum(~(be_p <= 1.0)))
    check["plain_info_nonzero"] = int((info != 0).sum())
    if args.check:
        # pivots, info and factors of a sample against the oracle, bit for bit
        from oracle import pyoracle
        k = 32
        LUo, ipo, info_o = pyoracle.batch_dgetrf(A0[:k].cpu().numpy())
        check["plain_oracle_sample"] = k
        check["plain_ipiv_equal"] = bool(np.array_equal(ipiv.numpy()[:k], ipo))
        check["plain_info_equal"] = bool(np.array_equal(info[:k].cpu().numpy(), info_o))
        check["plain_lu_bitexact"] = bool(np.array_equal(copies[-1][0][:k].cpu().numpy(), LUo))
    del copies, Xp
    # the plain factorization against both of its bounds
    fbytes = count * getrf_bytes(n)
    res["plain"]["roofline"] = dict(
        {"kernel": "getrf_mfma_kernel<256>", "fp64_tflops": count * getrf_flops(n) / (fact_ms * 1e-3) / 1e12,
         "fp64_frac": count * getrf_flops(n) / (fact_ms * 1e-3) / FP64_PEAK,
         "algorithmic_GBps": fbytes / (fact_ms * 1e-3) / 1e9},
        **measured_traffic("cond256_mixed_vs_plain", "getrf_mfma_kernel", fbytes))
    if "traffic" in res["plain"]["roofline"]:
        res["plain"]["roofline"]["traffic_GBps"] = res["plain"]["roofline"]["traffic"] / (fact_ms * 1e-3) / 1e9

    X = torch.zeros_like(B0)
    px = ma.ptr_array(X, count, n * nrhs)
    it = torch.zeros(count, dtype=torch.int32, device="cuda")
    mix = Timer(torch)

    def make_m(i):
        # A is read-only unless a matrix falls back to FP64 (then it holds
        # the LU factors), so every step gets its own copy as well
        A = A0.clone()
        return (A, B0, ma.ptr_array(A, count, n * n), ma.ptr_array(B0, count, n * nrhs))

    def step_mixed(c, timer):
        A, B, pa, pb = c
        e = mix.start() if timer else None
        L.magma_dsgesv_batched(n, nrhs, pa.data_ptr(), n, ipiv.ptr, pb.data_ptr(), n, px.data_ptr(), n,
                               info.data_ptr(), it.data_ptr(), count, q.handle)
        if timer:
            mix.stop(e)

    wall_m, _, _ = run_fixed(args, torch, ma, L, q, make_m, step_mixed, count, count, "matrices")
    mixed_ms = wall_m / args.steps * 1e3
    mixed_steps = np.asarray(mix.ms())
    iters = it.cpu().numpy()
    res["mixed"] = dict(solves_per_s=count / (mixed_ms * 1e-3), ms_per_step=mixed_ms,
                        kernel_ms=float(np.mean(mixed_steps)),
                        step_p50_ms=float(np.percentile(mixed_steps, 50)),
                        step_p99_ms=float(np.percentile(mixed_steps, 99)),
                        sweeps_mean=float(np.mean(np.maximum(iters, 0))),
                        sweeps_max=int(np.max(iters)),
                        fallbacks=int(np.sum(iters < 0)))
    res["mixed_over_plain"] = res["mixed"]["solves_per_s"] / res["plain"]["solves_per_s"]
    be_m = gpu_backward_errors(torch, A0, X, B0)
    check["mixed_backward_err_over_n_eps_max"] = float(be_m.max())
    check["mixed_backward_err_above_n_eps"] = int(np.sum(~(be_m <= 1.0)))
    check["backward_err_checked"] = f"all {count} solutions, FP64 residual on the device"
    if args.check:
        from oracle import pyoracle
        k = 32
        A_s = A0[:k].cpu().numpy(); B_s = B0[:k].cpu().numpy(); Xm = X[:k].cpu().numpy()
        check["mixed_oracle_residual_sample_max_over_n_eps"] = float(
            max(pyoracle.solve_error(A_s[s], Xm[s], B_s[s]) for s in range(k)))
    byt = count * (8.0 * n * n + 16.0 * n * nrhs + 8)
    roof = {"bound": "hbm", "achieved": byt / (res["mixed"]["kernel_ms"] * 1e-3) / 1e9,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "kernel": "mixed_kernel (FP32 factor + FP64 refinement; the dsgesv call's dominant launch)",
            "kernel_ms": res["mixed"]["kernel_ms"], "bytes_per_launch": byt}
    roof.update(measured_traffic("cond256_mixed_vs_plain", "mixed_kernel", byt))
    serve = None
    if args.serve_leg > 0:
        sargs = argparse.Namespace(**vars(args))
        sargs.serve_seconds = args.serve_leg
        sr = workload_serve(sargs, torch, ma, L, q)
        c = sr["config"]
        serve = {k: c[k] for k in ("arrival_rate_per_s", "seconds", "requests", "latency_p50_ms", "latency_p99_ms",
                                   "latency_max_ms", "fallbacks_fp64", "backward_err_over_n_eps_max",
                                   "reused_factorizations")}
        serve["solves_per_s"] = sr["value"]
        serve["workload"] = "serve_poisson (Poisson arrivals, n log-uniform [16,512], nrhs 1-16, 50% SPD, " \
                            "20% repeat right-hand sides, 64 GB budget)"
    return dict(value=res["mixed"]["solves_per_s"], unit="solves/s", ms=mixed_ms, dtype="f64",
                config={"workload": "cond256_mixed_vs_plain", "batch": count, "n": n, "nrhs": nrhs,
                        "inputs": "A = U diag(s) V', U,V from QR of N(0,1) (seed 3), cond log-uniform [1e2,1e6]",
                        "modes": res, "serve": serve},
                roofline=roof, check=check, step_ms=list(mixed_steps),
                cpu=dict(kind="dgesv", n=n, nrhs=nrhs, gen=lambda k: (A0[:k].cpu().numpy(), B0[:k].cpu().numpy())))


def workload_vbatch(args, torch, ma, L, q):
    from magma_amd import gen
    count = 50000
    sizes = gen.vbatch_sizes(2, count, 8, 256).astype(np.int32)
    lds = sizes.astype(np.int64)
    offs = np.concatenate([[0], np.cumsum(lds * lds)[:-1]])
    total = int((lds * lds).sum())
    dev = "cuda"
    Aflat0 = torch.empty(total, dtype=torch.float64, device=dev)
    Bflat = torch.empty(total, dtype=torch.float64, device=dev)
    Cflat = torch.empty(total, dtype=torch.float64, device=dev)
    # fill per size class (same generator: seed 2 for A, 102 for B, 202 for C)
    for nval in np.unique(sizes):
        idx = np.nonzero(sizes == nval)[0]
        nn = int(nval) * int(nval)
        for seed, buf in ((2, Aflat0), (102, Bflat), (202, Cflat)):
            chunk = max(1, (1 << 22) // nn)
            for c0 in range(0, len(idx), chunk):
                ids = idx[c0:c0 + chunk]
                mats = torch.from_numpy(ids.astype(np.int64)).to(dev).view(-1, 1)
                e = torch.arange(nn, dtype=torch.int64, device=dev).view(1, -1)
                vals = 2.0 * gen.u01_torch(seed, mats, e) - 1.0
                pos = torch.from_numpy(offs[ids]).to(dev).view(-1, 1) + e
                buf[pos.view(-1)] = vals.view(-1)
    d_n = ma.dev_ints(sizes)
    ipiv_off = np.concatenate([[0], np.cumsum(lds)[:-1]])
    ipiv = torch.zeros(int(lds.sum()), dtype=torch.int32, device=dev)
    pip = torch.from_numpy(ipiv.data_ptr() + ipiv_off * 4).to(dev)
    info = torch.zeros(count + 1, dtype=torch.int32, device=dev)
    pb = torch.from_numpy(Bflat.data_ptr() + offs * 8).to(dev)
    pc = torch.from_numpy(Cflat.data_ptr() + offs * 8).to(dev)
    lu_t, gemm_t = Timer(torch), Timer(torch)

    def make(i):
        A = Aflat0.clone()
        return (A, torch.from_numpy(A.data_ptr() + offs * 8).to(dev))

    def step(c, timer):
        A, pa = c
        e = lu_t.start() if timer else None
        L.magma_dgetrf_vbatched_max_nocheck(d_n.data_ptr(), d_n.data_ptr(), 256, pa.data_ptr(), d_n.data_ptr(),
                                            pip.data_ptr(), info.data_ptr(), count, q.handle)
        if timer:
            lu_t.stop(e)
            e = gemm_t.start()
        L.magmablas_dgemm_vbatched_max_nocheck(ma.MagmaNoTrans, ma.MagmaNoTrans, d_n.data_ptr(), d_n.data_ptr(),
                                               d_n.data_ptr(), 1.0, pa.data_ptr(), d_n.data_ptr(), pb.data_ptr(),
                                               d_n.data_ptr(), 1.0, pc.data_ptr(), d_n.data_ptr(), count, 256,
                                               256, 256, q.handle)
        if timer:
            gemm_t.stop(e)

    wall, _, _ = run_fixed(args, torch, ma, L, q, make, step, count, count, "matrices")
    ms = wall / args.steps * 1e3
    lu_ms = float(np.mean(lu_t.ms())); gemm_ms = float(np.mean(gemm_t.ms()))
    flops_lu = float(sum(getrf_flops(int(x)) for x in sizes))
    flops_gemm = float(np.sum(2.0 * lds ** 3))
    return dict(value=count / (ms * 1e-3), unit="factorizations+gemms/s", ms=ms, dtype="f64",
                config={"workload": "vbatch_lu_gemm", "batch": count, "n": "log-uniform [8,256] seed 2",
                        "lu_ms": lu_ms, "gemm_ms": gemm_ms,
                        "lu_tflops": flops_lu / (lu_ms * 1e-3) / 1e12,
                        "gemm_tflops": flops_gemm / (gemm_ms * 1e-3) / 1e12},
                roofline={"bound": "mfma", "achieved": flops_lu / (lu_ms * 1e-3) / 1e12,
                          "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s", "kernel": "getrf_vbatched (all classes)",
                          "kernel_ms": lu_ms, "flops_per_launch": flops_lu},
                check={}, step_ms=None, cpu=None)


def serve_trace(seed, rate, seconds, pool, nlo=16, nhi=512, reuse=0.2):
    """Poisson arrivals at `rate`/s for `seconds` s: per request (kind, n, nrhs,
    matrix id, factor key, reuses).  New requests draw a matrix from the pool
    and keep their factorization resident under their own key; a repeat
    request (probability `reuse`) takes a recent request's matrix and key with
    new right-hand sides."""
    from magma_amd import gen
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, int(rate * seconds * 1.1) + 16))
    arrivals = arrivals[arrivals < seconds]
    m = len(arrivals)
    pool_n = gen.vbatch_sizes(seed, pool, nlo, nhi)
    pool_kind = (gen.scalar_u(seed + 1, np.arange(pool)) < 0.5).astype(np.int32)     # 1 = SPD
    mat = rng.integers(0, pool, m)
    key = np.arange(m, dtype=np.int64)
    rep = rng.uniform(size=m) < reuse
    back = rng.integers(1, 2000, m)
    for i in np.nonzero(rep)[0]:
        j = i - back[i]
        if j >= 0:
            mat[i] = mat[j]
            key[i] = key[j]
        else:
            rep[i] = False
    nrhs = rng.integers(1, 17, m).astype(np.int32)
    return arrivals, mat, key, rep, nrhs, pool_n.astype(np.int32), pool_kind


def workload_serve(args, torch, ma, L, q):
    """Serving loop: Poisson arrivals, LU and SPD requests of varying n and
    nrhs, FP64 answers by mixed-precision refinement, 20% repeat requests,
    continuous batching under a 64 GB device budget (magma_amd_server_*)."""
    import ctypes
    from magma_amd import gen, lib
    rate, seconds = 20000.0, args.serve_seconds
    pool, ring = 4096, 32768
    arrivals, mat, key, rep, nrhs, pool_n, pool_kind = serve_trace(4, rate, seconds, pool)
    m = len(arrivals)
    log(f"serve trace: {m} requests over {seconds} s, building the matrix pool")
    # matrix pool on the device (column-major, ld = n)
    offs = np.concatenate([[0], np.cumsum(pool_n.astype(np.int64) ** 2)[:-1]])
    Apool = torch.empty(int((pool_n.astype(np.int64) ** 2).sum()), dtype=torch.float64, device="cuda")
    for k0 in range(0, pool, 256):
        for k in range(k0, min(pool, k0 + 256)):
            n = int(pool_n[k])
            if pool_kind[k]:
                G = gen.normal_torch(4, k, 1, n * n, "cuda").view(n, n)
                M = G @ G.T + n * torch.eye(n, dtype=torch.float64, device="cuda")
            else:
                M = gen.uniform_pm1_torch(4, k, 1, n * n, "cuda").view(n, n)
            Apool[int(offs[k]): int(offs[k]) + n * n] = M.reshape(-1)
    Bpool = gen.uniform_pm1_torch(104, 0, 1, 512 * 16 * 4096, "cuda").view(4096, 16, 512)
    Xring = torch.zeros((ring, 16, 512), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    log("pool ready")
    base_a, base_b, base_x = Apool.data_ptr(), Bpool.data_ptr(), Xring.data_ptr()
    idx = np.arange(m)
    rq = np.zeros(m, dtype=np.dtype([("kind", "i4"), ("n", "i4"), ("nrhs", "i4"), ("p0", "i4"),
                                      ("dA", "u8"), ("ldda", "i4"), ("p1", "i4"), ("dB", "u8"), ("lddb", "i4"),
                                      ("p2", "i4"), ("dX", "u8"), ("lddx", "i4"), ("p3", "i4"),
                                      ("key", "i8")], align=True))
    assert rq.dtype.itemsize == ctypes.sizeof(lib.Request)
    n_req = pool_n[mat]
    rq["kind"] = pool_kind[mat]
    rq["n"] = n_req
    rq["nrhs"] = nrhs
    rq["dA"] = base_a + offs[mat].astype(np.uint64) * 8
    rq["ldda"] = n_req
    rq["dB"] = base_b + (idx % 4096).astype(np.uint64) * (16 * 512 * 8)
    rq["lddb"] = 512
    rq["dX"] = base_x + (idx % ring).astype(np.uint64) * (16 * 512 * 8)
    rq["lddx"] = 512
    rq["key"] = key
    out = (lib.Completion * m)()
    h = ctypes.c_void_p()
    budget = 64 << 30
    assert L.magma_amd_server_create(0, budget, 512, ctypes.byref(h)) == 0
    try:
        # warm-up: a short trace to create the streams' first kernels
        warm = min(m, 2000)
        warm_out = (lib.Completion * warm)()
        wrq = rq[:warm].copy()
        wrq["key"] = -1                    # nothing kept resident from the warm-up
        warr = np.ascontiguousarray(arrivals[:warm] - arrivals[0])
        assert L.magma_amd_server_run_trace(h, wrq.ctypes.data, warr.ctypes.data, warm, warm_out) == 0
        log(f"warm-up done; replaying {m} requests")
        t0 = time.perf_counter()
        rc = L.magma_amd_server_run_trace(h, rq.ctypes.data, np.ascontiguousarray(arrivals).ctypes.data, m, out)
        wall = time.perf_counter() - t0
        assert rc == 0, rc
        st = lib.ServerStats()
        L.magma_amd_server_stats(h, ctypes.byref(st))
    finally:
        L.magma_amd_server_destroy(h)
    res = np.frombuffer(out, dtype=np.dtype([("ticket", "i8"), ("info", "i4"), ("iter", "i4"), ("reused", "i4"),
                                             ("pad", "i4"), ("berr", "f8"), ("ts", "f8"), ("tst", "f8"),
                                             ("td", "f8")], align=True), count=m)
    lat = (res["td"] - res["ts"]) * 1e3
    span = float(res["td"].max() - res["ts"].min())
    be_ratio = res["berr"] / (n_req * EPS)
    conv = res["iter"] >= 0
    log(f"served {m} in {wall:.1f} s: p50 {np.percentile(lat, 50):.3f} ms p99 {np.percentile(lat, 99):.3f} ms")
    return dict(value=m / span, unit="solves/s", ms=span * 1e3 / m, dtype="f64 (mixed f32/f16 factor)",
                config={"workload": "serve_poisson", "arrival_rate_per_s": rate, "seconds": seconds,
                        "requests": m, "n": "log-uniform [16,512]", "nrhs": "uniform [1,16]",
                        "spd_fraction": float(np.mean(pool_kind[mat])), "repeat_fraction": float(np.mean(rep)),
                        "budget_bytes": budget, "latency_p50_ms": float(np.percentile(lat, 50)),
                        "latency_p99_ms": float(np.percentile(lat, 99)), "latency_max_ms": float(lat.max()),
                        "reused_factorizations": int(res["reused"].sum()),
                        "fallbacks_fp64": int((~conv).sum()), "info_nonzero": int((res["info"] != 0).sum()),
                        "backward_err_over_n_eps_max": float(be_ratio[conv].max()) if conv.any() else None,
                        "backward_err_over_n_eps_p99": float(np.percentile(be_ratio[conv], 99)) if conv.any() else None,
                        "server": {"batches": int(st.batches), "max_batch": int(st.max_batch_seen),
                                   "evictions": int(st.evictions), "bytes_peak": int(st.bytes_peak),
                                   "resident_at_end": int(st.resident)}},
                roofline={"bound": "latency", "achieved": m / span, "peak": rate, "unit": "requests/s",
                          "kernel": "mixed_kernel (per request workgroup)", "note": "offered load bound"},
                check={}, step_ms=None, cpu=None)


def workload_blas(args, torch, ma, L, q):
    """magmablas_dtrsm_batched (Left, Lower, NoTrans, NonUnit, 16 right-hand
    sides), magmablas_dtrsv_batched (Upper, NoTrans) and
    magmablas_dgemv_batched (NoTrans) on 20,000 matrices of order 64, each
    timed on its own.  One step runs the three calls."""
    count, n, nrhs = 20000, 64, 16
    rng = np.random.default_rng(5)
    A0 = rng.uniform(-1, 1, (count, n, n)) / 8.0
    idx = np.arange(n)
    A0[:, idx, idx] = rng.uniform(1, 2, (count, n))
    A0 = torch.from_numpy(A0).cuda()
    B0 = torch.from_numpy(rng.uniform(-1, 1, (count, nrhs, n))).cuda()
    x0 = torch.from_numpy(rng.uniform(-1, 1, (count, n))).cuda()
    pa = ma.ptr_array(A0, count, n * n)
    timers = {k: Timer(torch) for k in ("trsm", "trsv", "gemv")}

    def make(i):
        B = B0.clone(); x = x0.clone(); y = torch.zeros_like(x0)
        return (B, x, y, ma.ptr_array(B, count, n * nrhs), ma.ptr_array(x, count, n), ma.ptr_array(y, count, n))

    def step(c, timer):
        B, x, y, pb, px, py = c
        for name in ("trsm", "trsv", "gemv"):
            e = timers[name].start() if timer else None
            if name == "trsm":
                L.magmablas_dtrsm_batched(ma.MagmaLeft, ma.MagmaLower, ma.MagmaNoTrans, ma.MagmaNonUnit, n, nrhs,
                                          1.0, pa.data_ptr(), n, pb.data_ptr(), n, count, q.handle)
            elif name == "trsv":
                L.magmablas_dtrsv_batched(ma.MagmaUpper, ma.MagmaNoTrans, ma.MagmaNonUnit, n, pa.data_ptr(), n,
                                          px.data_ptr(), 1, count, q.handle)
            else:
                L.magmablas_dgemv_batched(ma.MagmaNoTrans, n, n, 1.0, pa.data_ptr(), n, px.data_ptr(), 1, 0.0,
                                          py.data_ptr(), 1, count, q.handle)
            if timer:
                timers[name].stop(e)

    wall, step_ms, copies = run_fixed(args, torch, ma, L, q, make, step, count, count, "matrices")
    check = {}
    if args.check:
        from oracle import pyoracle
        A_h = A0[:16].cpu().numpy(); B_h = B0[:16].cpu().numpy(); x_h = x0[:16].cpu().numpy()
        B_g = copies[-1][0][:16].cpu().numpy(); x_g = copies[-1][1][:16].cpu().numpy()
        y_g = copies[-1][2][:16].cpu().numpy()
        ok = True
        for s in range(16):
            ok &= np.array_equal(B_g[s], pyoracle.dtrsm(b"L", b"L", b"N", b"N", n, nrhs, 1.0, A_h[s], B_h[s]))
            xs = pyoracle.dtrsv(b"U", b"N", b"N", n, A_h[s], x_h[s])
            ok &= np.array_equal(x_g[s], xs)
            ok &= np.array_equal(y_g[s], pyoracle.dgemv(b"N", n, n, 1.0, A_h[s], xs, 0.0, np.zeros(n)))
        check = {"sample": 16, "bitwise_equal_to_oracle": bool(ok)}
    ms = wall / args.steps * 1e3
    kms = {k: float(np.mean(t.ms())) for k, t in timers.items()}
    tri = 8.0 * n * (n + 1) / 2
    byts = {"trsm": count * (tri + 16.0 * n * nrhs), "trsv": count * (tri + 16.0 * n),
            "gemv": count * (8.0 * n * n + 8.0 * n + 8.0 * n)}
    rates = {k: byts[k] / (kms[k] * 1e-3) / 1e9 for k in kms}
    return dict(value=count / (ms * 1e-3), unit="matrices/s (trsm + trsv + gemv)", ms=ms, dtype="f64",
                config={"workload": "blas_batched", "batch": count, "n": n, "nrhs": nrhs,
                        "apis": ["magmablas_dtrsm_batched", "magmablas_dtrsv_batched", "magmablas_dgemv_batched"],
                        "kernel_ms": kms, "algorithmic_GBps": rates,
                        "inputs": "A U[-1,1]/8 with diagonal U[1,2] (seed 5)"},
                roofline={"bound": "hbm", "achieved": rates["trsm"], "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                          "kernel": "trsm_kernel<L, L, N>", "kernel_ms": kms["trsm"], "bytes_per_launch": byts["trsm"]},
                check=check, step_ms=step_ms, cpu=None)


def workload_extras(args, torch, ma, L, q):
    """magma_dgeqrf_batched, magma_dgetri_outofplace_batched (from the getrf
    factors, computed before the timed region), magma_dgesv_nopiv_batched and
    magma_dgesv_rbt_batched (one right-hand side) on 20,000 matrices of order
    64, each timed on its own; one step runs the four calls."""
    count, n = 20000, 64
    rng = np.random.default_rng(6)
    A0 = rng.uniform(-1, 1, (count, n, n))
    A0[:, np.arange(n), np.arange(n)] += n           # no-pivot LU stays accurate
    A0 = torch.from_numpy(A0).cuda()
    B0 = torch.from_numpy(rng.uniform(-1, 1, (count, 1, n))).cuda()
    LU = A0.clone()
    p_lu = ma.ptr_array(LU, count, n * n)
    ipiv = torch.zeros((count, n), dtype=torch.int32, device="cuda")
    p_ipiv = ma.ptr_array(ipiv, count, n)
    info = torch.zeros(count, dtype=torch.int32, device="cuda")
    L.magma_dgetrf_batched(n, n, p_lu.data_ptr(), n, p_ipiv.data_ptr(), info.data_ptr(), count, q.handle)
    torch.cuda.synchronize()
    names = ("geqrf", "getri", "gesv_nopiv", "gesv_rbt")
    timers = {k: Timer(torch) for k in names}

    def make(i):
        Aq = A0.clone(); tau = torch.zeros((count, n), dtype=torch.float64, device="cuda")
        inv = torch.zeros_like(A0)
        An = A0.clone(); Bn = B0.clone(); Ar = A0.clone(); Br = B0.clone()
        pp = [ma.ptr_array(t, count, t[0].numel()) for t in (Aq, tau, inv, An, Bn, Ar, Br)]
        return (Aq, tau, inv, An, Bn, Ar, Br), pp

    def step(c, timer):
        _, (pq, pt, pinv, pan, pbn, par, pbr) = c
        for name in names:
            e = timers[name].start() if timer else None
            if name == "geqrf":
                L.magma_dgeqrf_batched(n, n, pq.data_ptr(), n, pt.data_ptr(), info.data_ptr(), count, q.handle)
            elif name == "getri":
                L.magma_dgetri_outofplace_batched(n, p_lu.data_ptr(), n, p_ipiv.data_ptr(), pinv.data_ptr(), n,
                                                  info.data_ptr(), count, q.handle)
            elif name == "gesv_nopiv":
                L.magma_dgesv_nopiv_batched(n, 1, pan.data_ptr(), n, pbn.data_ptr(), n, info.data_ptr(), count,
                                            q.handle)
            else:
                L.magma_dgesv_rbt_batched(n, 1, par.data_ptr(), n, pbr.data_ptr(), n, info.data_ptr(), count,
                                          q.handle)
            if timer:
                timers[name].stop(e)

    wall, step_ms, copies = run_fixed(args, torch, ma, L, q, make, step, count, count, "matrices")
    check = {}
    if args.check:
        from oracle import pyoracle
        (Aq, tau, inv, An, Bn, Ar, Br), _ = copies[-1]
        A_h = A0[:8].cpu().numpy(); B_h = B0[:8].cpu().numpy()
        ok = True
        worst = 0.0
        for s in range(8):
            e1, e2 = pyoracle.qr_errors(A_h[s], Aq[s].cpu().numpy(), tau[s].cpu().numpy(), n, n)
            worst = max(worst, e1 / pyoracle.EPS, e2 / pyoracle.EPS)
            _, Xo, _ = pyoracle.dgesv_rbt(A_h[s], B_h[s])
            ok &= np.array_equal(Br[s].cpu().numpy(), Xo)
            LUo, info_o = pyoracle.dgetrf_nopiv(A_h[s])
            ok &= np.array_equal(An[s].cpu().numpy(), LUo)
        check = {"sample": 8, "qr_max_residual_over_eps": float(worst), "nopiv_rbt_bitwise_equal_to_oracle": bool(ok)}
    ms = wall / args.steps * 1e3
    kms = {k: float(np.mean(t.ms())) for k, t in timers.items()}
    # algorithmic bytes per matrix: the operands read and written once
    byts = {"geqrf": count * (16.0 * n * n + 8.0 * n), "getri": count * (8.0 * n * n + 4.0 * n + 8.0 * n * n),
            "gesv_nopiv": count * (16.0 * n * n + 16.0 * n), "gesv_rbt": count * (16.0 * n * n + 16.0 * n)}
    rates = {k: byts[k] / (kms[k] * 1e-3) / 1e9 for k in kms}
    flops = {"geqrf": count * (4.0 / 3.0) * n ** 3}
    return dict(value=count / (ms * 1e-3), unit="matrices/s (geqrf + getri + gesv_nopiv + gesv_rbt)", ms=ms,
                dtype="f64",
                config={"workload": "extras_batched", "batch": count, "n": n, "nrhs": 1, "kernel_ms": kms,
                        "algorithmic_GBps": rates,
                        "geqrf_tflops": flops["geqrf"] / (kms["geqrf"] * 1e-3) / 1e12,
                        "inputs": "A U[-1,1] + 64 I (seed 6), b U[-1,1]"},
                roofline={"bound": "hbm", "achieved": rates["geqrf"], "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                          "kernel": "geqrf_kernel", "kernel_ms": kms["geqrf"], "bytes_per_launch": byts["geqrf"]},
                check=check, step_ms=step_ms, cpu=None)


WORKLOADS = {
    "cond256_mixed_vs_plain": workload_cond256,
    "extras_batched": workload_extras,
    "blas_batched": workload_blas,
    "lu32_gesv": workload_lu32,
    "chol64_posv": workload_chol64,
    "vbatch_lu_gemm": workload_vbatch,
    "serve_poisson": workload_serve,
}


def cpu_baseline(spec, seconds):
    """The reference's CPU path on a bounded sample of the same workload:
    LAPACK dgetrf + dgetrs (dpotrf + dpotrs) called once per matrix in an
    OpenMP loop over the batch with LAPACK single-threaded, as MAGMA's testers
    run it (testing/testing_zgesv_batched.cpp:159-186).  The LAPACK is the
    host's OpenBLAS when one is found (then it is `value`); the oracle's
    netlib-order C restatement, under the same OpenMP loop, is reported beside
    it (and is `value` when no LAPACK is found)."""
    if spec is None:
        return None
    from oracle import pyoracle
    threads = cpu_threads()
    kind, n, nrhs = spec["kind"], spec["n"], spec["nrhs"]
    fn = pyoracle.batch_dgesv if kind == "dgesv" else pyoracle.batch_dposv
    lfn = pyoracle.lapack_batch_dgesv if kind == "dgesv" else pyoracle.lapack_batch_dposv
    k = max(threads, 8)
    A, B = spec["gen"](k)
    t0 = time.perf_counter(); fn(A, B, nthreads=threads); dt = time.perf_counter() - t0
    # at most 20000 distinct matrices in host memory; the sample is repeated
    # until it holds about `seconds` of CPU work
    k2 = int(min(max(k, k * seconds / max(dt, 1e-6)), 20000))
    A, B = spec["gen"](k2)
    t0 = time.perf_counter(); fn(A, B, nthreads=threads); dt = time.perf_counter() - t0
    reps = max(1, int(np.ceil(seconds / max(dt, 1e-6))) - 1)
    log(f"cpu baseline: {k2} matrices x {reps + 1} passes")
    t0 = time.perf_counter()
    for _ in range(reps):
        fn(A, B, nthreads=threads)
    dt = (dt + time.perf_counter() - t0) / (reps + 1)
    port = {"value": k2 / dt, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{k2} matrices of the workload (n={n}, nrhs={nrhs}) x {reps + 1} passes, oracle {kind} "
                      f"(netlib-order C restatement) under OpenMP, {threads} threads, {dt * (reps + 1):.1f} s"}
    cfg = pyoracle.lapack()
    if not cfg:
        return port
    lreps = max(2, reps + 1)
    t0 = time.perf_counter()
    for _ in range(lreps):
        lfn(A, B, nthreads=threads)
    dl = (time.perf_counter() - t0) / lreps
    routine = "dgetrf + dgetrs" if kind == "dgesv" else "dpotrf + dpotrs"
    return {"value": k2 / dl, "unit": "solves/s", "cores": threads, "kind": "reference",
            "path": f"the reference's CPU path: LAPACK {routine} per matrix, OpenMP loop over the batch "
                    f"(testing/testing_zgesv_batched.cpp:159-186), LAPACK = {cfg.split()[0]} single-threaded",
            "library": cfg,
            "sample": f"{k2} matrices of the workload (n={n}, nrhs={nrhs}) x {lreps} passes, {threads} threads, "
                      f"{dl * lreps:.1f} s",
            "port": port}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cond256_mixed_vs_plain", choices=sorted(WORKLOADS))
    ap.add_argument("--batch", type=int, default=0, help="override the batch count (cond256 only)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--serve-seconds", type=float, default=30.0, help="trace length of serve_poisson")
    ap.add_argument("--serve-leg", type=float, default=5.0,
                    help="seconds of the serving trace replayed after the cond256 workload (0: none); its "
                         "p50/p99 latency go into config.serve")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one 1 GiB device-to-device copy (the reference kernel for scaling "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip magma_amd_autotune and use the untuned kernel variants")
    ap.add_argument("--tune-table", default="",
                    help="load the variant table from this file if it exists (no tuning run), "
                         "else tune and save it there")
    ap.add_argument("--pin", action="append", default=[], metavar="FAMILY:CLASS:VARIANT",
                    help="after tuning (or instead of it, with --no-autotune), fix the kernel variant of one "
                         "size class, e.g. mixed:5:4 (families getrf, mixed, gemm; classes 0-9 as in "
                         "magma_amd_tune_set); repeatable, for reproducible profiling runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", dest="check", action="store_false")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: the batched path has no cross-GPU step "
              "(every matrix is independent); run one process per GPU and add the rates.", file=sys.stderr)
        sys.exit(2)

    import torch
    import magma_amd as ma
    L = ma.load()
    q = ma.Queue(0)
    L.magma_queue_set_async(q.handle, 1)      # back-to-back submission; synchronize around the timed region

    if args.calibrate_copy:
        # an elementwise kernel of known traffic: reads 1 GiB, writes 1 GiB
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
        log("calibration kernel: 1 GiB read + 1 GiB written")
    tuned = None
    if args.autotune:
        t0 = time.perf_counter()
        if args.tune_table and os.path.exists(args.tune_table):
            rc = L.magma_amd_tune_load(args.tune_table.encode())
            if rc != 0:
                raise RuntimeError(f"magma_amd_tune_load({args.tune_table}) failed: {rc}")
        else:
            rc = L.magma_amd_autotune(q.handle, 0)
            if rc != 0:
                raise RuntimeError(f"magma_amd_autotune failed: {rc}")
            if args.tune_table and L.magma_amd_tune_save(args.tune_table.encode()) != 0:
                raise RuntimeError(f"magma_amd_tune_save({args.tune_table}) failed")
        tuned = {name: [L.magma_amd_tune_get(fam, c) for c in range(10)]
                 for fam, name in enumerate(("getrf", "mixed", "gemm"))}
        log(f"autotune {time.perf_counter() - t0:.1f} s: {tuned}")
    families = {"getrf": 0, "mixed": 1, "gemm": 2}
    for pin in args.pin:
        fam, cls, var = pin.split(":")
        if L.magma_amd_tune_set(families[fam], int(cls), int(var)) != 0:
            raise RuntimeError(f"--pin {pin}: no such family, class or variant")
    if args.pin:
        tuned = {name: [L.magma_amd_tune_get(fam, c) for c in range(10)]
                 for fam, name in enumerate(("getrf", "mixed", "gemm"))}
        log(f"pinned variants: {tuned}")
    log(f"workload {args.workload}")
    r = WORKLOADS[args.workload](args, torch, ma, L, q)
    roof = r["roofline"]
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof.setdefault("traffic", None)
    cpu = None if args.no_cpu_baseline else cpu_baseline(r.get("cpu"), args.cpu_seconds)
    lat = {}
    if r.get("step_ms"):
        sm = np.asarray(r["step_ms"])
        lat = {"step_p50_ms": float(np.percentile(sm, 50)), "step_p99_ms": float(np.percentile(sm, 99))}
    out = {
        "metric": METRIC,
        "value": r["value"],
        "unit": r["unit"],
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": r["dtype"],
        "data": "synthetic (deterministic splitmix64 generator, inputs resident in HBM)",
        "config": dict(r["config"], parallelism="single GPU, batch-parallel", tuned_variants=tuned, **lat),
        "roofline": roof,
        "cpu_baseline": cpu,
        "check": r["check"],
    }
    print(json.dumps(out), flush=True)
    q.close()


if __name__ == "__main__":
    main()
