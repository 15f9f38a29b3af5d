#!/usr/bin/env python3
"""Benchmark: matrices solved per second by the compact batched SPD solve on one MI355X.

Workload (the configuration the headline metric is quoted on): dposv_compact,
n = 32, nrhs = 4, 131072 SPD matrices A = Q diag(lambda) Q^T (Q from the QR of an
N(0,1) matrix, log10(lambda) uniform on [0, k], k drawn per matrix from {2, 4, 6}),
B ~ N(0,1), packed in the compact layout with V = 64, resident in HBM.

A step is one call over the whole batch. Two paths are timed:
  plain  dposv_compact   (FP64 factor + solve, A and B overwritten; each step
                          runs on its own pre-filled copy, so no restore sits
                          inside the timed region)
  mixed  dsposv_compact  (FP32 factor, FP64 residual refinement, FP64 fallback;
                          A and B are inputs, X the output), when the library
                          exports it.
value = the headline path's matrices/s counting only matrices whose backward
error max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf) is
<= 1e-14 (computed in FP64 on the GPU after the timed steps).

Also reported, in the same JSON line:
  dgels_n32    the least-squares half of the metric: dgels_compact (plain
               FP64) against dsgels_compact (FP32 QR preconditioner, FP64
               residuals) at m=64, n=32, nrhs=2, 131072 problems, counting
               problems with ||A^T r||_2 / (||A||_F ||r||_2) <= 1e-14;
  configs      the other configurations (K1 dposv n=8, K2 dgels/dgeqrf/dormqr
               16x8, K3 sysvnp and trsm 24x24, the mixed dssysvnp at K3's
               shape, device pack/unpack of 32x32 FP64), one call each;
  uneven_gels  dgels_batch / sgels_batch / dsgels_batch (FP32 batch QR +
               FP64 refinement + FP64 fallback) over 500000 least-squares
               problems of uneven sizes (one recorded graph launch each),
               against one dgels_compact launch per distinct size (all 16104
               of them), with the LAPACK-test LS backward error of a sample;
  roofline     of the dominant kernel (algorithmic bytes per launch / its
               average duration from HIP events on the launch stream, against
               MI355X's 8 TB/s; traffic from the rocprofv3 counters recorded in
               profiles/);
  cpu_baseline the reference's own CPU path (oracle/_ref, the reference's
               src/cbk.cpp compiled from its sources) timed on this host in
               a child process with OMP_PROC_BIND=close OMP_PLACES=cores, on
               every physical core this process may use and on 1 thread:
               dposv n=32 nrhs=4 (the headline shape), K1 exactly as
               configured (dposv n=8 nrhs=1, 65536 matrices, V=8) and dgels
               64x32 nrhs=2; lscpu's sockets x cores x threads recorded.
"""

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, NRHS, NM, V = 32, 4, 131072, 64
HBM_PEAK_GBS = 8000.0
ETA_MAX = 1e-14


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nm", type=int, default=NM)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-batch gels leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the K1-K3 configuration leg")
    ap.add_argument("--uneven-nprob", type=int, default=K5_NPROB)
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the tuners (their launches share kernel names and grids with the "
                         "timed legs, which a profile has to keep apart); the uneven leg then "
                         "plans with the default tuning")
    ap.add_argument("--cpu-leg", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a plain 1 GiB device-to-device copy (the known-size kernel "
                         "that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(nm, seed=2026):
    """dense FP64 A (nm, n, n) and B (nm, n, nrhs) on the host (numpy): A = Q
    diag(lambda) Q^T with Q from the QR of an N(0,1) matrix (sign-fixed so
    diag(R) > 0), log10(lambda) ~ U[0, k], k from {2, 4, 6} per matrix,
    symmetrised; B ~ N(0, 1)"""
    import numpy as np
    r = np.random.default_rng(seed)
    A = np.empty((nm, N, N))
    chunk = 16384
    for s in range(0, nm, chunk):
        e = min(nm, s + chunk)
        G = r.standard_normal((e - s, N, N))
        Q, Rr = np.linalg.qr(G)
        Q = Q * np.sign(np.diagonal(Rr, axis1=1, axis2=2))[:, None, :]
        k = np.array([2.0, 4.0, 6.0])[r.integers(0, 3, e - s)]
        lam = 10.0 ** (r.uniform(0, 1, (e - s, N)) * k[:, None])
        M = (Q * lam[:, None, :]) @ np.swapaxes(Q, 1, 2)
        A[s:e] = 0.5 * (M + np.swapaxes(M, 1, 2))
    B = r.standard_normal((nm, N, NRHS))
    return A, B


def pack(torch, X, V):
    """dense (nm, rows, cols) -> compact column-major, ld = rows: (ng, cols, rows, V)"""
    nm, r, c = X.shape
    ng = (nm + V - 1) // V
    full = torch.zeros(ng * V, r, c, dtype=X.dtype, device=X.device)
    full[:nm] = X
    for d in range(min(r, c)):
        full[nm:, d, d] = 1.0
    return full.view(ng, V, r, c).permute(0, 3, 2, 1).contiguous()


def unpack(torch, P, nm):
    ng, c, r, V = P.shape
    return P.permute(0, 3, 2, 1).reshape(ng * V, r, c)[:nm]


def backward_error(torch, A, X, B):
    """per matrix max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf), FP64"""
    R = B - torch.matmul(A, X)
    an = A.abs().sum(dim=2).amax(dim=1)
    num = R.abs().amax(dim=1)
    den = an[:, None] * X.abs().amax(dim=1) + B.abs().amax(dim=1)
    return (num / den).amax(dim=1)


def load_traffic():
    """HBM bytes per launch of the two kernels, measured with rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE on this benchmark and scaled by the same counters on
    a plain copy of known size (profiles/kernel_traffic.json), or None"""
    path = os.path.join(ROOT, "profiles", "kernel_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def host_topology():
    """sockets, cores per socket and threads per core (lscpu), and the CPUs
    this process may run on"""
    import subprocess
    topo = {"sockets": None, "cores_per_socket": None, "threads_per_core": None}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                topo[keys[k.strip()]] = int(v.strip())
        for line in out.splitlines():
            if line.startswith("Model name"):
                topo["model"] = line.partition(":")[2].strip()
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    topo["affinity_cpus"] = len(os.sched_getaffinity(0))
    return topo


def cpu_leg_main():
    """Child process of the CPU baseline (OpenMP pinned by the environment the
    parent sets): times the reference's own kernels (oracle/_ref, its
    src/cbk.cpp compiled from its sources; the oracle's restatement if that
    build is absent) on bounded samples of the workloads and prints one JSON
    line. Best of 3 after a warm-up, inputs restored outside the timed calls
    (the reference's best_time, examples/bench_util.hpp:250-267)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cbkt
    path = cbkt.ref_path()
    kind = "reference" if path else "port"
    lib, prefix = (ctypes.CDLL(path), "") if path else (ctypes.CDLL(cbkt.ORACLE_SO), "oracle_")
    name = os.path.basename(path) if path else "oracle port"
    posv = getattr(lib, prefix + "dposv_compact")
    posv.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                     ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    posv.restype = ctypes.c_int
    gels = getattr(lib, prefix + "dgels_compact")
    gels.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                     ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                     ctypes.c_int, ctypes.c_int]
    gels.restype = ctypes.c_int
    r = np.random.default_rng(2030)

    def best_of(call, restore):
        best = float("inf")
        for rep in range(4):
            restore()
            t0 = time.perf_counter()
            assert call() == 0
            dt = time.perf_counter() - t0
            if rep > 0:
                best = min(best, dt)
        return best

    def spd(nm, n, kmax=(2, 4, 6)):
        G = r.standard_normal((nm, n, n))
        Q, Rr = np.linalg.qr(G)
        lam = 10.0 ** (r.uniform(0, 1, (nm, n)) * np.array(kmax)[r.integers(0, len(kmax), nm)][:, None])
        return (Q * lam[:, None, :]) @ np.swapaxes(Q, 1, 2)

    out = {}
    # K4 shape: dposv n=32 nrhs=4 FP64, V=8, 131072 matrices (the headline's distribution)
    nm, n, q, Vc = 131072, 32, 4, 8
    A = spd(nm, n)
    B = r.standard_normal((nm, n, q))
    ap0, bp0 = cbkt.pack(A, Vc, fill=0.0), cbkt.pack(B, Vc, pad_diag=0.0, fill=0.0)
    ap, bp = ap0.copy(), bp0.copy()

    def restore():
        ap[...] = ap0
        bp[...] = bp0
    t = best_of(lambda: posv(b"C", b"L", n, q, ap.ctypes.data, n, bp.ctypes.data, n, Vc, nm), restore)
    X = cbkt.unpack(bp, nm, n, q, Vc)
    eta = float(np.max(np.abs(B - A @ X).max(axis=1) / (np.abs(A).sum(axis=2).max(axis=1)[:, None] *
                                                        np.abs(X).max(axis=1) + np.abs(B).max(axis=1))))
    out["K4_dposv_n32"] = {"value": nm / t, "unit": "matrices/s", "seconds": t,
                           "sample": f"dposv_compact n=32 nrhs=4 FP64 V=8, {nm} matrices, {name}; "
                                     f"max backward error {eta:.1e}"}
    del A, B, ap0, bp0, ap, bp
    # K1 exactly as configured: dposv n=8 nrhs=1, 65536 matrices, V=8
    nm, n, q = 65536, 8, 1
    G = r.standard_normal((nm, n, n))
    A = G @ np.swapaxes(G, 1, 2) + 8 * np.eye(n)
    B = r.standard_normal((nm, n, q))
    ap0, bp0 = cbkt.pack(A, Vc, fill=0.0), cbkt.pack(B, Vc, pad_diag=0.0, fill=0.0)
    ap, bp = ap0.copy(), bp0.copy()
    t = best_of(lambda: posv(b"C", b"L", n, q, ap.ctypes.data, n, bp.ctypes.data, n, Vc, nm), restore)
    out["K1_dposv_n8"] = {"value": nm / t, "unit": "matrices/s", "seconds": t,
                          "sample": f"dposv_compact n=8 nrhs=1 FP64 V=8, {nm} matrices, {name}"}
    # dgels n=32: 64 x 32, nrhs=2, U(-1,1), 131072 problems, V=8
    nm, m, n, q = 131072, 64, 32, 2
    A = r.uniform(-1, 1, (nm, m, n))
    B = r.uniform(-1, 1, (nm, m, q))
    ap0, bp0 = cbkt.pack(A, Vc, fill=0.0), cbkt.pack(B, Vc, pad_diag=0.0, fill=0.0)
    ap, bp = ap0.copy(), bp0.copy()
    tau = np.zeros(((nm + Vc - 1) // Vc) * Vc * n)
    t = best_of(lambda: gels(b"C", b"N", m, n, q, ap.ctypes.data, m, bp.ctypes.data, m,
                             tau.ctypes.data, Vc, nm), restore)
    out["dgels_64x32"] = {"value": nm / t, "unit": "matrices/s", "seconds": t,
                          "sample": f"dgels_compact 64x32 nrhs=2 FP64 V=8, {nm} problems, {name}"}
    print(json.dumps({"kind": kind, "legs": out}), flush=True)


def cpu_baseline():
    """the reference's CPU path on this host: a child process per thread
    count (all the physical cores this process may use, then 1), with
    OMP_PROC_BIND=close and OMP_PLACES=cores"""
    import subprocess
    topo = host_topology()
    tpc = topo.get("threads_per_core") or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    phys = max(1, topo["affinity_cpus"] // tpc)
    threads = min(phys, share) if share > 0 else phys
    runs = {}
    for nt in (threads, 1):
        env = dict(os.environ, OMP_NUM_THREADS=str(nt), OMP_PROC_BIND="close", OMP_PLACES="cores")
        p = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg"], env=env,
                           capture_output=True, text=True, timeout=900)
        if p.returncode != 0:
            raise RuntimeError(f"CPU leg failed ({p.returncode}): {p.stderr[-2000:]}")
        runs[nt] = json.loads(p.stdout.strip().splitlines()[-1])
    full, one = runs[threads], runs[1]
    head = full["legs"]["K4_dposv_n32"]
    legs = {}
    for k, v in full["legs"].items():
        legs[k] = dict(v, threads=threads)
        legs[k + "_1thread"] = dict(one["legs"][k], threads=1)
    return {"value": head["value"], "unit": "matrices/s", "cores": threads, "kind": full["kind"],
            "sample": head["sample"] + f"; {threads} threads, OMP_PROC_BIND=close OMP_PLACES=cores",
            "host": dict(topo, omp_proc_bind="close", omp_places="cores"), "legs": legs}


GELS_M, GELS_N, GELS_NRHS, GELS_NM = 64, 32, 2, 131072


def gels_bench(torch, K, W, nm=GELS_NM):
    """dgels_compact (plain FP64) against dsgels_compact (FP32 QR preconditioner,
    FP64 residuals) on m=64, n=32, nrhs=2 least-squares problems, A and B
    entries ~ U(-1,1), compact V=64. Counted: problems whose backward error
    ||A^T r||_2 / (||A||_F ||r||_2) is <= 1e-14 (FP64, on the GPU)."""
    import cbk_amd
    m, n, q = GELS_M, GELS_N, GELS_NRHS
    g = torch.Generator(device="cuda")
    g.manual_seed(2028)
    Ad = torch.rand(nm, m, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    Bd = torch.rand(nm, m, q, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    A0, B0 = pack(torch, Ad, V), pack(torch, Bd, V)
    T = torch.empty(((nm + V - 1) // V) * n * V, dtype=torch.float64, device="cuda")
    X = torch.zeros(((nm + V - 1) // V), q, n, V, dtype=torch.float64, device="cuda")
    it = torch.zeros(nm, dtype=torch.int32, device="cuda")
    info = torch.zeros(nm, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def eta(Xs):
        R = Bd - torch.matmul(Ad, Xs)
        return (torch.matmul(Ad.transpose(1, 2).contiguous(), R).norm(dim=(1, 2)) /
                (Ad.norm(dim=(1, 2)) * R.norm(dim=(1, 2))))

    A, B = A0.clone(), B0.clone()
    tp = []
    for s in range(W + K):
        A.copy_(A0)
        B.copy_(B0)
        ev[0].record(stream)
        assert cbk_amd.call("d", "gels", "C", "N", m, n, q, A, m, B, m, T, V, nm) == 0
        ev[1].record(stream)
        torch.cuda.synchronize()
        if s >= W:
            tp.append(ev[0].elapsed_time(ev[1]) / 1e3)
    e_plain = eta(unpack(torch, B, nm)[:, :n])
    tm = []
    for s in range(W + K):
        A.copy_(A0)
        ev[0].record(stream)
        assert cbk_amd.call_ext("dsgels_compact", "C", "N", m, n, q, A, m, B0, m, X, n, it, info,
                                V, nm) == 0
        ev[1].record(stream)
        torch.cuda.synchronize()
        if s >= W:
            tm.append(ev[0].elapsed_time(ev[1]) / 1e3)
    e_mixed = eta(unpack(torch, X, nm))
    t_p, t_m = sum(tp) / K, sum(tm) / K
    ok_p, ok_m = int((e_plain <= ETA_MAX).sum().item()), int((e_mixed <= ETA_MAX).sum().item())
    vals, cnts = torch.unique(it.cpu(), return_counts=True)
    out = {"workload": f"dgels_compact / dsgels_compact m={m} n={n} nrhs={q}, {nm} problems, "
                       "U(-1,1), compact V=64",
           "plain_fp64": {"matrices_per_s": ok_p / t_p, "kernel_ms": t_p * 1e3, "ok_matrices": ok_p,
                          "max_backward_error": e_plain.max().item()},
           "mixed": {"matrices_per_s": ok_m / t_m, "kernel_ms": t_m * 1e3, "ok_matrices": ok_m,
                     "max_backward_error": e_mixed.max().item(),
                     "iter_histogram": {int(a): int(b) for a, b in zip(vals, cnts)}},
           "backward_error": "||A^T r||_2 / (||A||_F ||r||_2) <= 1e-14"}
    out["mixed_over_plain"] = out["mixed"]["matrices_per_s"] / out["plain_fp64"]["matrices_per_s"]
    log(f"dgels n=32: plain {t_p * 1e3:.2f} ms, mixed {t_m * 1e3:.2f} ms, "
        f"mixed/plain {out['mixed_over_plain']:.2f}, eta {e_plain.max().item():.1e} / {e_mixed.max().item():.1e}")
    del Ad, Bd, A0, B0, A, B, T, X
    torch.cuda.empty_cache()
    return out


def configs_bench(torch):
    """the other configurations of the hot path, each timed as one call over
    inputs resident in HBM (best of 3 after a warm-up; inputs restored
    outside the timed region), as matrices/s and algorithmic GB/s:
      K1  dposv n=8 nrhs=1, 2^21 matrices (the CPU config's shape, V=64 on the GPU)
      K2  dgels 16x8 nrhs=2, 2^20; dgeqrf and dormqr (L, T) of the same shape
      K3  dsysvnp / ssysvnp n=24 nrhs=24 (off-diagonals N(0,1), diagonal +-26),
          dtrsm / strsm L L N U 24x24 with 24 right-hand sides, 262144 matrices"""
    import cbk_amd
    g = torch.Generator(device="cuda")
    g.manual_seed(2029)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    stream = torch.cuda.current_stream()
    out = {}

    def timeit(f, setup):
        ts = []
        for i in range(4):
            setup()
            torch.cuda.synchronize()
            ev[0].record(stream)
            assert f() == 0
            ev[1].record(stream)
            torch.cuda.synchronize()
            if i:
                ts.append(ev[0].elapsed_time(ev[1]) / 1e3)
        return min(ts)

    def rec(key, t, nm, bpm, what):
        out[key] = {"what": what, "us": t * 1e6, "matrices_per_s": nm / t,
                    "algorithmic_gbs": bpm * nm / t / 1e9, "frac_of_8tbs": bpm * nm / t / 8e12}
        log(f"{key}: {t * 1e6:.1f} us, {nm / t:.3e} matrices/s, {bpm * nm / t / 1e9:.0f} GB/s")

    nm, n = 1 << 21, 8
    G = torch.randn(nm, n, n, dtype=torch.float64, device="cuda", generator=g)
    A0 = pack(torch, G @ G.transpose(1, 2) + 8 * torch.eye(n, dtype=torch.float64, device="cuda"), V)
    del G
    B0 = pack(torch, torch.randn(nm, n, 1, dtype=torch.float64, device="cuda", generator=g), V)
    A, B = A0.clone(), B0.clone()
    t = timeit(lambda: cbk_amd.call("d", "posv", "C", "L", n, 1, A, n, B, n, V, nm),
               lambda: (A.copy_(A0), B.copy_(B0)))
    rec("K1_dposv_n8", t, nm, 8 * (n * (n + 1) + 2 * n), "dposv_compact n=8 nrhs=1, A=GG^T+8I, 2^21, V=64")
    del A0, B0, A, B

    nm, m, n, q = 1 << 20, 16, 8, 2
    A0 = pack(torch, torch.rand(nm, m, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1, V)
    B0 = pack(torch, torch.rand(nm, m, q, dtype=torch.float64, device="cuda", generator=g) * 2 - 1, V)
    T = torch.empty(nm * n, dtype=torch.float64, device="cuda")
    A, B = A0.clone(), B0.clone()
    t = timeit(lambda: cbk_amd.call("d", "gels", "C", "N", m, n, q, A, m, B, m, T, V, nm),
               lambda: (A.copy_(A0), B.copy_(B0)))
    rec("K2_dgels_16x8", t, nm, 8 * (2 * m * n + n + 2 * m * q), "dgels_compact 16x8 nrhs=2, U(-1,1), 2^20")
    t = timeit(lambda: cbk_amd.call("d", "geqrf", "C", m, n, A, m, T, V, nm), lambda: A.copy_(A0))
    rec("K2_dgeqrf_16x8", t, nm, 8 * (2 * m * n + n), "dgeqrf_compact 16x8")
    t = timeit(lambda: cbk_amd.call("d", "ormqr", "C", "L", "T", m, q, n, A, m, T, B, m, V, nm),
               lambda: B.copy_(B0))
    # the reflector bodies below the diagonal are read, tau, and C read and written
    rec("K2_dormqr_16x2_k8", t, nm, 8 * (m * n - n * (n + 1) // 2 + n + 2 * m * q),
        "dormqr_compact L T, C 16x2, k=8")
    del A0, B0, A, B, T

    nm, n, q = 262144, 24, 24
    for p, dt, es in (("d", torch.float64, 8), ("s", torch.float32, 4)):
        M = torch.tril(torch.randn(nm, n, n, dtype=torch.float64, device="cuda", generator=g), -1)
        sign = torch.where(torch.rand(nm, n, device="cuda", generator=g) < 0.5, -26.0, 26.0).double()
        M = M + M.transpose(1, 2) + torch.diag_embed(sign)
        A0 = pack(torch, M.to(dt), V)
        del M
        B0 = pack(torch, torch.randn(nm, n, q, dtype=torch.float64, device="cuda", generator=g).to(dt), V)
        A, B = A0.clone(), B0.clone()
        t = timeit(lambda: cbk_amd.call(p, "sysvnp", "C", "L", n, q, A, n, B, n, V, nm),
                   lambda: (A.copy_(A0), B.copy_(B0)))
        rec(f"K3_{p}sysvnp_n24", t, nm, es * (n * (n + 1) + 2 * n * q), f"{p}sysvnp_compact n=24 nrhs=24")
        if p == "d" and hasattr(cbk_amd.lib(), "dssysvnp_compact"):
            # mixed: FP32 LDL^T, FP64 refinement; A and B read, X written
            X = torch.empty_like(B0)
            it = torch.empty(nm, dtype=torch.int32, device="cuda")
            inf = torch.empty_like(it)
            t = timeit(lambda: cbk_amd.call_ext("dssysvnp_compact", "C", "L", n, q, A, n, B0, n, X, n,
                                                None, 0, it, inf, V, nm), lambda: A.copy_(A0))
            rec("K3_dssysvnp_n24", t, nm, 8 * (n * (n + 1) // 2 + 2 * n * q) + 8,
                "dssysvnp_compact n=24 nrhs=24 (mixed)")
            out["K3_dssysvnp_n24"]["refined"] = int((it >= 0).sum().item())
            del X, it, inf
        t = timeit(lambda: cbk_amd.call(p, "trsm", "C", "L", "L", "N", "U", n, q, 1.0, A0, n, B, n, V, nm),
                   lambda: B.copy_(B0))
        rec(f"K3_{p}trsm_24x24", t, nm, es * (n * (n - 1) // 2 + 2 * n * q), f"{p}trsm_compact L L N U 24x24 nrhs=24")
        # side R: X A = B with A upper (the side-L kernels on B's transposed view)
        t = timeit(lambda: cbk_amd.call(p, "trsm", "C", "R", "U", "N", "U", n, q, 1.0, A0, n, B, n, V, nm),
                   lambda: B.copy_(B0))
        rec(f"K3_{p}trsm_R_24x24", t, nm, es * (n * (n - 1) // 2 + 2 * n * q),
            f"{p}trsm_compact R U N U 24x24, 24 rows of B")
        del A0, B0, A, B
    torch.cuda.empty_cache()

    # dense <-> compact on the device: 131072 FP64 32 x 32 column-major
    # matrices, V = 64 (each value read once and written once)
    nm, n = 131072, 32
    lib = cbk_amd.lib()
    if hasattr(lib, "cbk_dpack_compact"):
        D = torch.randn(nm * n * n, dtype=torch.float64, device="cuda", generator=g)
        P = torch.empty_like(D)
        t = timeit(lambda: cbk_amd.call_ext("cbk_dpack_compact", "C", n, n, D, "C", n, n * n, P, n,
                                            1.0, V, nm), lambda: None)
        rec("pack_d32", t, nm, 16 * n * n, "cbk_dpack_compact 32x32 FP64, 131072 matrices, V=64")
        t = timeit(lambda: cbk_amd.call_ext("cbk_dunpack_compact", "C", n, n, P, n, D, "C", n, n * n,
                                            V, nm), lambda: None)
        rec("unpack_d32", t, nm, 16 * n * n, "cbk_dunpack_compact 32x32 FP64")
        del D, P
        torch.cuda.empty_cache()
    return out


K5_NPROB = 500000


def uneven_inputs(nprob, seed=2027):
    """the uneven least-squares batch: m ~ U[4,64], n ~ U[2,m], nrhs ~ U[1,8],
    A and B entries ~ U(-1,1); each problem column-major with lda = ldb = m,
    packed back to back. Returns (descriptor array, A pool, B pool) on the host."""
    import numpy as np
    import cbk_amd
    r = np.random.default_rng(seed)
    m = r.integers(4, 65, nprob)
    n = 2 + (r.random(nprob) * (m - 1)).astype(np.int64)
    nrhs = r.integers(1, 9, nprob)
    asz, bsz = m * n, m * nrhs
    desc = np.zeros(nprob, cbk_amd.problem_dtype())
    desc["m"], desc["n"], desc["nrhs"], desc["lda"], desc["ldb"] = m, n, nrhs, m, m
    desc["a_off"] = np.cumsum(asz) - asz
    desc["b_off"] = np.cumsum(bsz) - bsz
    A = r.uniform(-1, 1, int(asz.sum()))
    B = r.uniform(-1, 1, int(bsz.sum()))
    return desc, A, B


def ls_backward_error(desc, Ah, Bh, Xpool, sample):
    """the largest least-squares backward-error ratio of LAPACK's own tests
    (dqrt17) over the sampled problems and their columns:
    ||A^T r||_2 / (||A||_F (||A||_F ||x||_2 + ||b||_2)), r = b - A x; an FP64
    Householder solve gives eps-level values"""
    import numpy as np
    worst = 0.0
    for k in sample:
        mk, nk, qk = int(desc["m"][k]), int(desc["n"][k]), int(desc["nrhs"][k])
        a0, b0 = int(desc["a_off"][k]), int(desc["b_off"][k])
        A = Ah[a0:a0 + mk * nk].reshape(nk, mk).T
        B = Bh[b0:b0 + mk * qk].reshape(qk, mk).T
        X = Xpool[b0:b0 + mk * qk].reshape(qk, mk).T[:nk]
        R = B - A @ X
        an = np.linalg.norm(A)
        eta = np.linalg.norm(A.T @ R, axis=0) / (an * (an * np.linalg.norm(X, axis=0) +
                                                       np.linalg.norm(B, axis=0)))
        worst = max(worst, float(eta.max()))
    return worst


TUNE_NM = 98304


def compact_tuner_bench():
    """cbk_autotune_compact at the configurations' shapes (98304 synthetic
    matrices each, a count no timed leg uses, so that profiles keep the
    tuner's dispatches apart): the kernel family it installs, and at K4's
    shape also the interleave width it picks among 16, 32 and 64. Runs after
    every timed leg (the families it installs would otherwise steer them)."""
    import cbk_amd
    names = {v: k for k, v in cbk_amd.FAMILY.items()}
    shapes = [("K1_dposv_n8", "d", "posv", 8, 8, 1, "L", "N"),
              ("K2_dgels_16x8", "d", "gels", 16, 8, 2, "L", "N"),
              ("K2_dgeqrf_16x8", "d", "geqrf", 16, 8, 0, "L", "N"),
              ("K3_dsysvnp_n24", "d", "sysvnp", 24, 24, 24, "L", "N"),
              ("K3_ssysvnp_n24", "s", "sysvnp", 24, 24, 24, "L", "N"),
              ("K3_dtrsm_24x24", "d", "trsm", 24, 24, 0, "L", "N"),
              ("K3_strsm_24x24", "s", "trsm", 24, 24, 0, "L", "N"),
              ("K4_dposv_n32", "d", "posv", 32, 32, 4, "L", "N")]
    out = {}
    for name, p, routine, m, n, nrhs, uplo, trans in shapes:
        t0 = time.perf_counter()
        st, fam, v = cbk_amd.autotune_compact(p, routine, m, n, nrhs, 64, nm=TUNE_NM, uplo=uplo,
                                              trans=trans)
        out[name] = {"status": st, "family": names.get(fam, fam), "V": v,
                     "seconds": time.perf_counter() - t0}
    t0 = time.perf_counter()
    st, fam, v = cbk_amd.autotune_compact("d", "posv", 32, 32, 4, 0, nm=TUNE_NM)
    out["K4_dposv_n32_pick_V"] = {"status": st, "family": names.get(fam, fam), "V": v,
                                  "seconds": time.perf_counter() - t0}
    log(f"compact tuner: {out}")
    return out


def uneven_bench(torch, K, W, nprob=K5_NPROB):
    """dgels_batch / sgels_batch over the uneven batch (one graph launch per
    step, inputs re-copied outside the timed region), against the same problems
    solved by dgels_compact once per distinct (m, n, nrhs), each size packed
    compact with V = 64, as a caller without a batch API would"""
    import numpy as np
    import cbk_amd
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    log(f"generating the uneven batch ({nprob} problems) ...")
    desc, Ah, Bh = uneven_inputs(nprob)
    m, n, q = desc["m"], desc["n"], desc["nrhs"]
    sample = np.random.default_rng(7).choice(nprob, 4000, replace=False)
    nbytes = {"D": 8, "S": 4}
    out = {"workload": f"gels, {nprob} problems, m~U[4,64], n~U[2,m], nrhs~U[1,8], U(-1,1), "
                       "column-major per problem",
           "nprob": nprob}
    for prec, dt in (("D", np.float64), ("S", np.float32)):
        A0 = torch.from_numpy(Ah.astype(dt)).cuda()
        B0 = torch.from_numpy(Bh.astype(dt)).cuda()
        A, B = A0.clone(), B0.clone()
        plan = cbk_amd.GelsBatchPlan(prec, "N", desc)
        work = torch.empty(max(plan.worksize, 1), dtype=torch.uint8, device="cuda")
        info = torch.empty(nprob, dtype=torch.int32, device="cuda")
        ts = []
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        stream = torch.cuda.current_stream()
        for s in range(W + K):
            A.copy_(A0)
            B.copy_(B0)
            e0.record(stream)
            assert plan.run(A, B, info, work) == 0, cbk_amd.last_error()
            e1.record(stream)
            torch.cuda.synchronize()
            if s >= W:
                ts.append(e0.elapsed_time(e1) / 1e3)
        t = sum(ts) / len(ts)
        traffic = 2 * nbytes[prec] * int((m * n + m * q).sum())
        res = {"ms": t * 1e3, "problems_per_s": nprob / t, "hbm_gbs": traffic / t / 1e9,
               "bytes": traffic, "classes": plan.classes(),
               "nonzero_info": int((info != 0).sum().item())}
        # a sample against the oracle's batch gels on the host
        if prec == "D":
            import batchcase
            ns = 2000
            idx = np.random.default_rng(5).choice(nprob, ns, replace=False)
            Ag, Bg = A.cpu().numpy(), B.cpu().numpy()
            worst = 0.0
            sd, sa, sb, As, Bs = [], [], [], [], []
            for k in idx:
                mk, nk, qk = int(m[k]), int(n[k]), int(q[k])
                a0, b0 = int(desc["a_off"][k]), int(desc["b_off"][k])
                As.append(Ah[a0:a0 + mk * nk].reshape(nk, mk).T.copy())
                Bs.append(Bh[b0:b0 + mk * qk].reshape(qk, mk).T.copy())
            sdesc, sap, sbp, _, _ = batchcase.make(np.random.default_rng(0),
                                                   [(int(m[k]), int(n[k]), int(q[k])) for k in idx], pad=(0,))
            for j, k in enumerate(idx):
                mk, nk, qk, lda, ldb, a0, b0 = sdesc[j]
                sap[a0:a0 + mk * nk] = As[j].T.reshape(-1)
                sbp[b0:b0 + mk * qk] = Bs[j].T.reshape(-1)
            assert batchcase.oracle_batch("d", "N", sdesc, sap, sbp) == 0
            for j, k in enumerate(idx):
                mk, nk, qk = int(m[k]), int(n[k]), int(q[k])
                b0 = int(desc["b_off"][k])
                Xg = Bg[b0:b0 + mk * qk].reshape(qk, mk).T[:nk]
                Xo = batchcase.get_b(sbp, sdesc[j])[:nk]
                worst = max(worst, float(np.abs(Xg - Xo).max() / max(np.abs(Xo).max(), 1e-300)))
            res["sample_vs_oracle"] = {"problems": ns, "max_rel_diff_x": worst}
        if prec == "D":
            res["sample_ls_backward_error"] = ls_backward_error(desc, Ah, Bh, B.cpu().numpy(), sample)
        out["fp64" if prec == "D" else "fp32"] = res
        log(f"uneven gels {prec}: {t * 1e3:.2f} ms/step, {nprob / t:.3e} problems/s, "
            f"{traffic / t / 1e9:.0f} GB/s, classes {plan.classes()}")
        del A0, B0, A, B, work, plan
        torch.cuda.empty_cache()

    # the mixed form: FP32 batch QR, FP64 refinement, FP64 fallback, one graph
    if hasattr(cbk_amd.lib(), "dsgels_batch"):
        A0 = torch.from_numpy(Ah).cuda()
        B0 = torch.from_numpy(Bh).cuda()
        A = A0.clone()
        X = torch.empty_like(B0)
        plan = cbk_amd.GelsBatchPlan("M", "N", desc)
        work = torch.empty(max(plan.worksize, 1), dtype=torch.uint8, device="cuda")
        it = torch.empty(nprob, dtype=torch.int32, device="cuda")
        info = torch.empty(nprob, dtype=torch.int32, device="cuda")
        ts = []
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        stream = torch.cuda.current_stream()
        for s_ in range(W + K):
            A.copy_(A0)
            e0.record(stream)
            assert plan.run_mixed(A, B0, X, it, info, work) == 0, cbk_amd.last_error()
            e1.record(stream)
            torch.cuda.synchronize()
            if s_ >= W:
                ts.append(e0.elapsed_time(e1) / 1e3)
        t = sum(ts) / len(ts)
        itc = it.cpu().numpy()
        vals, cnts = np.unique(itc, return_counts=True)
        # A and B read, X written (refined problems leave A untouched)
        traffic = 8 * int((m * n + 2 * m * q).sum())
        res = {"ms": t * 1e3, "problems_per_s": nprob / t, "hbm_gbs": traffic / t / 1e9,
               "bytes": traffic, "iter_histogram": {str(int(v)): int(c) for v, c in zip(vals, cnts)},
               "fp64_fallback": int((itc < 0).sum()), "nonzero_info": int((info != 0).sum().item()),
               "workspace_bytes": plan.worksize,
               "sample_ls_backward_error": ls_backward_error(desc, Ah, Bh, X.cpu().numpy(), sample)}
        out["mixed"] = res
        out["mixed_over_fp64"] = out["fp64"]["ms"] / res["ms"]
        log(f"uneven dsgels_batch: {t * 1e3:.2f} ms/step, {nprob / t:.3e} problems/s, "
            f"FP64 fallbacks {res['fp64_fallback']}, iters {res['iter_histogram']}, "
            f"LS backward error (sample) {res['sample_ls_backward_error']:.2e}")
        del A0, B0, A, X, work, plan
        torch.cuda.empty_cache()

    # per-size compact launches, FP64
    log("per-size compact launches ...")
    key = (m * 100 + n) * 10 + q
    order = np.argsort(key, kind="stable")
    ks = key[order]
    bounds = np.flatnonzero(np.diff(ks)) + 1
    starts = np.concatenate([[0], bounds])
    ends = np.concatenate([bounds, [nprob]])
    V = 64
    ntotal = len(starts)
    # every distinct size, one dgels_compact launch each (the groups padded to
    # whole V = 64 groups with identity matrices, as the compact API needs)
    pick = np.arange(ntotal)
    groups, apos, bpos = [], 0, 0
    for s0, s1 in zip(starts[pick], ends[pick]):
        ids = order[s0:s1]
        mk, nk, qk = int(m[ids[0]]), int(n[ids[0]]), int(q[ids[0]])
        ng = (len(ids) + V - 1) // V
        groups.append((mk, nk, qk, len(ids), ids, apos, bpos))
        apos += ng * V * mk * nk
        bpos += ng * V * mk * qk
    Apk = np.zeros(apos)
    Bpk = np.zeros(bpos)
    tau_max = 0
    for mk, nk, qk, cnt, ids, a0, b0 in groups:
        ng = (cnt + V - 1) // V
        ga = Ah[desc["a_off"][ids][:, None] + np.arange(mk * nk)]          # (cnt, n*m) col-major
        gb = Bh[desc["b_off"][ids][:, None] + np.arange(mk * qk)]
        fa = np.zeros((ng * V, mk * nk))
        fa[:cnt] = ga
        eye = np.zeros((mk, nk))
        eye[np.arange(nk), np.arange(nk)] = 1.0
        fa[cnt:] = eye.T.reshape(-1)
        fb = np.zeros((ng * V, mk * qk))
        fb[:cnt] = gb
        Apk[a0:a0 + ng * V * mk * nk] = fa.reshape(ng, V, mk * nk).transpose(0, 2, 1).reshape(-1)
        Bpk[b0:b0 + ng * V * mk * qk] = fb.reshape(ng, V, mk * qk).transpose(0, 2, 1).reshape(-1)
        tau_max = max(tau_max, ng * V * nk)
    dA = torch.from_numpy(Apk).cuda()
    dB = torch.from_numpy(Bpk).cuda()
    dT = torch.empty(tau_max, dtype=torch.float64, device="cuda")
    fn_, _ = cbk_amd.fn("d", "gels")
    base_a, base_b, base_t = dA.data_ptr(), dB.data_ptr(), dT.data_ptr()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    stream = torch.cuda.current_stream()
    e0.record(stream)
    for mk, nk, qk, cnt, ids, a0, b0 in groups:
        st = fn_(b"C", b"N", mk, nk, qk, ctypes.c_void_p(base_a + 8 * a0), mk,
                 ctypes.c_void_p(base_b + 8 * b0), mk, ctypes.c_void_p(base_t), V, cnt)
        assert st == 0
    e1.record(stream)
    torch.cuda.synchronize()
    tp = e0.elapsed_time(e1) / 1e3
    nsamp = sum(g[3] for g in groups)
    out["per_size_fp64"] = {"ms": tp * 1e3, "problems_per_s": nsamp / tp,
                            "launches": len(groups), "problems": nsamp}
    out["batch_over_per_size_fp64"] = out["fp64"]["problems_per_s"] / out["per_size_fp64"]["problems_per_s"]
    if "mixed" in out:
        out["mixed_over_per_size_fp64"] = out["mixed"]["problems_per_s"] / out["per_size_fp64"]["problems_per_s"]
    log(f"per-size dgels_compact: all {ntotal} sizes, {nsamp} problems, {tp * 1e3:.1f} ms")
    del dA, dB, dT
    torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.cpu_leg:
        cpu_leg_main()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for {args.gpus}). "
              "Matrices are independent, so N GPUs would each run this kernel on nm/N matrices.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import cbk_amd

    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    cbk_amd.set_stream(stream)
    nm = args.nm
    ng = (nm + V - 1) // V
    K, W = args.steps, args.warmup

    log(f"generating {nm} SPD matrices n={N} (cond up to 1e6) ...")
    A_host, B_host = make_inputs(nm)
    A = torch.from_numpy(A_host).cuda()
    B = torch.from_numpy(B_host).cuda()
    Ap, Bp = pack(torch, A, V), pack(torch, B, V)

    # ---- plain FP64 dposv: one pre-filled copy per step
    nsets = K + W
    Aw = [Ap.clone() for _ in range(nsets)]
    Bw = [Bp.clone() for _ in range(nsets)]
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(nsets)]
    for s in range(W):
        assert cbk_amd.call("d", "posv", "C", "L", N, NRHS, Aw[s], N, Bw[s], N, V, nm) == 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(W, nsets):
        ev[s][0].record(stream)
        assert cbk_amd.call("d", "posv", "C", "L", N, NRHS, Aw[s], N, Bw[s], N, V, nm) == 0
        ev[s][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kt = [ev[s][0].elapsed_time(ev[s][1]) / 1e3 for s in range(W, nsets)]
    avg_plain = sum(kt) / len(kt)
    X = unpack(torch, Bw[-1], nm)
    eta = backward_error(torch, A, X, B)
    ok_plain = int((eta <= ETA_MAX).sum().item())
    log(f"plain dposv: avg kernel {avg_plain * 1e6:.1f} us, wall/step {wall / K * 1e6:.1f} us, "
        f"max eta {eta.max().item():.2e}, ok {ok_plain}/{nm}")
    del Aw, Bw
    torch.cuda.empty_cache()

    bytes_per_matrix = 8 * (N * (N + 1) + 2 * N * NRHS)
    plain_rate = ok_plain / (wall / K)

    # ---- mixed dsposv: A and B are inputs, X the output; one set serves all steps
    Xp = torch.zeros_like(Bp)
    iters = torch.zeros(nm, dtype=torch.int32, device="cuda")
    info = torch.zeros(nm, dtype=torch.int32, device="cuda")
    ws = cbk_amd.dsposv_worksize("C", "L", N, NRHS, N, N, N, V, nm)
    work = torch.empty(max(ws, 8), dtype=torch.uint8, device="cuda") if ws else None
    Am = Ap.clone()

    def mixed_step():
        return cbk_amd.call_ext("dsposv_compact", "C", "L", N, NRHS, Am, N, Bp, N, Xp, N, work, ws,
                                iters, info, V, nm)

    for s in range(W):
        assert mixed_step() == 0
    torch.cuda.synchronize()
    evm = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(K)]
    t0 = time.perf_counter()
    for s in range(K):
        evm[s][0].record(stream)
        assert mixed_step() == 0
        evm[s][1].record(stream)
    torch.cuda.synchronize()
    wall_m = time.perf_counter() - t0
    avg_mixed = sum(a.elapsed_time(b) for a, b in evm) / K / 1e3
    Xm = unpack(torch, Xp, nm)
    eta_m = backward_error(torch, A, Xm, B)
    ok_mixed = int((eta_m <= ETA_MAX).sum().item())
    it_cpu = iters.cpu()
    hist = {int(k): int(v) for k, v in zip(*torch.unique(it_cpu, return_counts=True))}
    mixed_rate = ok_mixed / (wall_m / K)
    log(f"mixed dsposv: avg kernel {avg_mixed * 1e6:.1f} us, wall/step {wall_m / K * 1e6:.1f} us, "
        f"max eta {eta_m.max().item():.2e}, ok {ok_mixed}/{nm}, iter histogram {hist}")
    mixed_bytes = 8 * (N * (N + 1) // 2 + 2 * N * NRHS)

    result = {
        "metric": "matrices solved/sec (dposv, dgels n=32) mixed vs FP64 at backward err<=1e-14",
        "value": mixed_rate,
        "unit": "matrices/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall_m / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factor + f64 refinement",
        "data": "synthetic: A = Q diag(10^(k*U[0,1])) Q^T, k in {2,4,6} per matrix; B ~ N(0,1); seeded",
        "config": {"workload": "dsposv_compact (mixed dposv) n=32 nrhs=4, 131072 matrices, compact V=64",
                   "nm": nm, "n": N, "nrhs": NRHS, "V": V},
        "mixed": {"matrices_per_s": mixed_rate, "kernel_us": avg_mixed * 1e6, "ok_matrices": ok_mixed,
                  "max_backward_error": eta_m.max().item(), "iter_histogram": hist,
                  "bytes_per_matrix": mixed_bytes,
                  "hbm_gbs": mixed_bytes * nm / avg_mixed / 1e9},
        "plain_fp64": {"matrices_per_s": plain_rate, "kernel_us": avg_plain * 1e6,
                       "ok_matrices": ok_plain, "max_backward_error": eta.max().item(),
                       "bytes_per_matrix": bytes_per_matrix,
                       "hbm_gbs": bytes_per_matrix * nm / avg_plain / 1e9},
        "mixed_over_plain": mixed_rate / plain_rate if plain_rate else None,
        "roofline": {
            "bound": "hbm",
            "achieved": mixed_bytes * nm / avg_mixed / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": mixed_bytes * nm / avg_mixed / 1e9 / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "dsposv_compact (mixed_posv_kernel<32,4>)",
            "bytes_per_matrix": mixed_bytes,
        },
    }
    traffic = load_traffic() or {}
    if "mixed_posv_kernel_dsposv_n32_nrhs4" in traffic:
        result["roofline"]["traffic"] = traffic["mixed_posv_kernel_dsposv_n32_nrhs4"]["bytes_per_launch"]
    if "quad_sym_kernel_dposv_n32_nrhs4" in traffic:
        result["plain_fp64"]["traffic_bytes_per_launch"] = traffic["quad_sym_kernel_dposv_n32_nrhs4"]["bytes_per_launch"]
    if args.calibrate:
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
    result["dgels_n32"] = gels_bench(torch, K=max(2, min(K, 10)), W=1)
    if not args.no_configs:
        result["configs"] = configs_bench(torch)
    if not args.no_uneven:
        # the uneven batch's class thresholds and grid sizes from the tuner
        tuned = None
        if not args.no_tune:
            log("batch tuner (cbk_autotune) ...")
            t0 = time.perf_counter()
            st = cbk_amd.autotune()
            tuned = cbk_amd.get_tuning()
            tuned["seconds"] = time.perf_counter() - t0
            tuned["status"] = st
            log(f"batch tuning: {tuned}")
        result["uneven_gels"] = uneven_bench(torch, K=max(2, min(K, 5)), W=1, nprob=args.uneven_nprob)
        result["uneven_gels"]["tuning"] = tuned if tuned else cbk_amd.get_tuning()
    if not args.no_configs and not args.no_tune:
        result["compact_tuner"] = compact_tuner_bench()
    if not args.no_cpu:
        log("CPU baseline (reference dposv_compact on the host) ...")
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
