"""Uneven gels batches for the tests: generation (each problem column-major
in a pool, with leading-dimension padding), the oracle's batch call, and the
per-problem checks against it."""

import ctypes

import numpy as np

import cbkt


class OracleProblem(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int), ("n", ctypes.c_int), ("nrhs", ctypes.c_int),
                ("lda", ctypes.c_int), ("ldb", ctypes.c_int),
                ("a_off", ctypes.c_longlong), ("b_off", ctypes.c_longlong)]


def make(r, shapes, dtype=np.float64, pad=(0, 2)):
    """shapes: list of (m, n, nrhs). Returns (desc, A pool, B pool, dense A list,
    dense B list); A ~ U(-1,1), B ~ U(-1,1), lda = m + pad, ldb = max(m,n) + pad"""
    desc, As, Bs = [], [], []
    ao = bo = 0
    for k, (m, n, nrhs) in enumerate(shapes):
        mx = max(m, n)
        pa = pad[k % len(pad)]
        lda, ldb = max(1, m) + pa, max(1, mx) + pa
        As.append(r.uniform(-1, 1, (m, n)).astype(dtype))
        Bs.append(r.uniform(-1, 1, (mx, nrhs)).astype(dtype))
        desc.append((m, n, nrhs, lda, ldb, ao, bo))
        ao += lda * max(n, 1) + 3
        bo += ldb * max(nrhs, 1) + 1
    ap = np.full(ao, 7.0, dtype)
    bp = np.full(bo, 7.0, dtype)
    for (m, n, nrhs, lda, ldb, a0, b0), A, B in zip(desc, As, Bs):
        for j in range(n):
            ap[a0 + j * lda: a0 + j * lda + m] = A[:, j]
        for j in range(nrhs):
            bp[b0 + j * ldb: b0 + j * ldb + B.shape[0]] = B[:, j]
    return desc, ap, bp, As, Bs


def get_a(ap, d):
    m, n, nrhs, lda, ldb, a0, b0 = d
    return np.stack([ap[a0 + j * lda: a0 + j * lda + m] for j in range(n)], axis=1) \
        if n else np.zeros((m, 0), ap.dtype)


def get_b(bp, d):
    m, n, nrhs, lda, ldb, a0, b0 = d
    mx = max(m, n)
    return np.stack([bp[b0 + j * ldb: b0 + j * ldb + mx] for j in range(nrhs)], axis=1) \
        if nrhs else np.zeros((mx, 0), bp.dtype)


def oracle_batch(p, trans, desc, ap, bp, info=None):
    arr = (OracleProblem * max(1, len(desc)))()
    for i, d in enumerate(desc):
        arr[i] = OracleProblem(*d)
    f = getattr(cbkt.oracle().lib, f"oracle_{p}gels_batch")
    f.restype = ctypes.c_int
    return f(ctypes.c_char(trans.encode()), ctypes.c_int(len(desc)), arr,
             ap.ctypes.data_as(ctypes.c_void_p), bp.ctypes.data_as(ctypes.c_void_p),
             None if info is None else info.ctypes.data_as(ctypes.c_void_p))


def uneven_shapes(r, nprob, mmax=64, rmax=8):
    """the K5 distribution: m ~ U[4, mmax], n ~ U[2, m], nrhs ~ U[1, rmax]"""
    out = []
    for _ in range(nprob):
        m = int(r.integers(4, mmax + 1))
        n = int(r.integers(2, m + 1))
        out.append((m, n, int(r.integers(1, rmax + 1))))
    return out


def check_ls(p, desc, As, Bs, ap_o, bp_o, ap_g, bp_g, factor=20.0):
    """GPU vs oracle per overdetermined problem, within the least-squares
    perturbation bound: R and reflectors to factor*m*eps*||A||, X to
    factor*m*eps*(kappa + kappa^2 ||r||/(||A|| ||x||))*||x||, Q^T B's tail to
    factor*m*eps*kappa*||b||"""
    eps = float(np.finfo(cbkt.DT[p]).eps)
    worst = 0.0
    for k, d in enumerate(desc):
        m, n, nrhs = d[:3]
        if nrhs == 0 or n == 0 or m < n:
            continue
        A = As[k].astype(np.float64)
        B0 = Bs[k].astype(np.float64)
        Fo, Fg = get_a(ap_o, d).astype(np.float64), get_a(ap_g, d).astype(np.float64)
        amax = np.abs(A).max()
        ef = np.abs(Fo - Fg).max() / (factor * m * eps * amax)
        s = np.linalg.svd(A, compute_uv=False)
        kappa = s[0] / s[-1]
        Xo = get_b(bp_o, d).astype(np.float64)
        Xg = get_b(bp_g, d).astype(np.float64)
        X = Xo[:n]
        r = B0 - A @ X
        xn = max(np.abs(X).max(), 1e-300)
        bound = factor * m * eps * (kappa + kappa ** 2 * np.linalg.norm(r) /
                                    (s[0] * max(np.linalg.norm(X), 1e-300)))
        ex = np.abs(Xg[:n] - X).max() / (bound * xn)
        er = 0.0
        if m > n:
            er = np.abs(Xg[n:] - Xo[n:]).max() / (factor * m * eps * kappa * max(np.abs(B0).max(), 1e-300))
        worst = max(worst, ef, ex, er)
        assert ef <= 1 and ex <= 1 and er <= 1, (k, d, ef, ex, er)
    return worst


def set_a(ap, d, A):
    m, n, nrhs, lda, ldb, a0, b0 = d
    for j in range(n):
        ap[a0 + j * lda: a0 + j * lda + m] = A[:, j]


def ill_conditioned(r, m, n, logk=12):
    """m x n, singular values 10^(-logk * U[0,1]) with 1 and 10^-logk present"""
    U = np.linalg.qr(r.standard_normal((m, n)))[0]
    W = np.linalg.qr(r.standard_normal((n, n)))[0]
    s = 10.0 ** (-r.uniform(0, 1, n) * logk)
    s[0], s[-1] = 1.0, 10.0 ** -logk
    return (U * s) @ W.T


def oracle_mixed(trans, desc, ap, bp):
    """oracle_dsgels_batch on copies: (A after, X, iter, info)"""
    arr = (OracleProblem * max(1, len(desc)))()
    for i, d in enumerate(desc):
        arr[i] = OracleProblem(*d)
    ap = ap.copy()
    xp = np.full_like(bp, 7.0)
    it = np.zeros(len(desc), np.int32)
    info = np.zeros(len(desc), np.int32)
    f = cbkt.oracle().lib.oracle_dsgels_batch
    f.restype = ctypes.c_int
    st = f(ctypes.c_char(trans.encode()), ctypes.c_int(len(desc)), arr,
           ap.ctypes.data_as(ctypes.c_void_p), bp.ctypes.data_as(ctypes.c_void_p),
           xp.ctypes.data_as(ctypes.c_void_p), it.ctypes.data_as(ctypes.c_void_p),
           info.ctypes.data_as(ctypes.c_void_p))
    assert st == 0
    return ap, xp, it, info


def mixed_mix(r, nuneven=200):
    """a batch for the mixed path: shapes it refines (every refinement bucket),
    problems it must hand to FP64 (cond 1e12), and shapes outside its range;
    returns (shapes, indices of the ill-conditioned, indices outside)"""
    shapes = [
        (16, 8, 2), (12, 5, 3), (32, 16, 5), (30, 30, 1),          # rows <= 32, n <= 16 / 32
        (64, 8, 8), (50, 13, 4), (64, 32, 2), (40, 31, 7),         # rows <= 64, n <= 16 / 32
        (64, 17, 8), (4, 2, 1), (2, 1, 1),
    ] + [s for s in uneven_shapes(r, nuneven) if s[1] <= 32]
    ill = list(range(len(shapes), len(shapes) + 4))
    shapes += [(24, 12, 2), (64, 30, 3), (32, 32, 6), (48, 20, 8)]
    outside = list(range(len(shapes), len(shapes) + 8))
    shapes += [(70, 10, 2), (30, 40, 2), (20, 5, 9), (5, 0, 2), (6, 3, 0), (80, 64, 4),
               (64, 33, 2), (40, 40, 8)]
    return shapes, ill, outside
