"""Pin the oracle's surfaces that have no reference counterpart, and the
reference algorithms themselves, against LAPACK (scipy.linalg.lapack) and
closed forms. A planted mistake in each (a wrong index, a dropped term) must
fail one of these checks.

  - info codes of potrf/posv/gels against LAPACK dpotrf / the first zero R(i,i);
  - posv X and potrf L against LAPACK dposv / dpotrf (the SPD factor is unique);
  - geqrf R and tau against LAPACK dgeqrf (same sign convention), gels X against
    LAPACK dgels;
  - sytrfnp against the closed form A = L D L^T of a constructed (L, D);
  - dsposv (LAPACK dsposv restated): X against dposv, iter >= 0 on well-
    conditioned matrices, iter < 0 plus the FP64 factor on cond ~ 1e12.
"""

import numpy as np
import scipy.linalg.lapack as lapack

import cbkt


def test_potrf_posv_match_lapack_and_info():
    r = cbkt.rng(11)
    nm, n, V = 21, 12, 8
    A = cbkt.gen_spd(r, nm, n)
    A[4, 6, 6] = -50.0   # leading minor 7 not positive definite
    B = r.standard_normal((nm, n, 3))
    ap, bp = cbkt.pack(A, V), cbkt.pack(B, V, pad_diag=0.0)
    info = np.zeros(nm, np.int32)
    assert cbkt.oracle().call("d", "posv", "C", "L", n, 3, ap, n, bp, n, V, nm, info=info) == 0
    L = cbkt.unpack(ap, nm, n, n, V)
    X = cbkt.unpack(bp, nm, n, 3, V)
    for k in range(nm):
        c, inf = lapack.dpotrf(A[k], lower=1)
        assert info[k] == inf
        if inf == 0:
            assert np.allclose(np.tril(L[k]), np.tril(c), rtol=0, atol=1e-12 * np.abs(c).max())
            x, inf2 = lapack.dpotrs(c, B[k], lower=1)
            assert np.allclose(X[k], x, rtol=1e-11, atol=1e-13)
        else:
            assert inf == 7
            assert np.array_equal(X[k], B[k])   # LAPACK dposv leaves B when dpotrf fails


def test_plants_are_caught():
    """the checks above are discriminating: a transposed factor or a sign flip fails"""
    r = cbkt.rng(2)
    A = cbkt.gen_spd(r, 1, 6)[0]
    c, _ = lapack.dpotrf(A, lower=1)
    assert not np.allclose(np.tril(c).T, np.tril(c))
    assert not np.allclose(-np.tril(c), np.tril(c))


def test_geqrf_gels_match_lapack():
    r = cbkt.rng(5)
    nm, m, n, V = 13, 9, 5, 4
    A = cbkt.gen_uniform(r, nm, m, n)
    B = r.uniform(-1, 1, (nm, m, 2))
    ap = cbkt.pack(A, V)
    tp = np.zeros(((nm + V - 1) // V) * n * V)
    assert cbkt.oracle().call("d", "geqrf", "C", m, n, ap, m, tp, V, nm) == 0
    F = cbkt.unpack(ap, nm, m, n, V)
    T = cbkt.unpack_vec(tp, nm, n, V)
    ap2, bp2 = cbkt.pack(A, V), cbkt.pack(B, V, pad_diag=0.0)
    tp2 = np.zeros_like(tp)
    info = np.zeros(nm, np.int32)
    assert cbkt.oracle().call("d", "gels", "C", "N", m, n, 2, ap2, m, bp2, m, tp2, V, nm, info=info) == 0
    X = cbkt.unpack(bp2, nm, m, 2, V)[:, :n]
    for k in range(nm):
        qr, tau, work, inf = lapack.dgeqrf(A[k])
        assert np.allclose(F[k], qr, atol=1e-13) and np.allclose(T[k], tau, atol=1e-14)
        lqr, x, inf2 = lapack.dgels(A[k], B[k])[:3]
        assert inf2 == 0 and np.allclose(X[k], x[:n], atol=1e-12)
    assert np.all(info == 0)


def test_gels_info_zero_column():
    r = cbkt.rng(6)
    nm, m, n, V = 5, 8, 4, 4
    A = cbkt.gen_uniform(r, nm, m, n)
    A[2, :, 1] = 0.0
    ap = cbkt.pack(A, V)
    bp = cbkt.pack(r.uniform(-1, 1, (nm, m, 1)), V, pad_diag=0.0)
    tp = np.zeros(((nm + V - 1) // V) * n * V)
    info = np.zeros(nm, np.int32)
    assert cbkt.oracle().call("d", "gels", "C", "N", m, n, 1, ap, m, bp, m, tp, V, nm, info=info) == 0
    assert info.tolist() == [0, 0, 2, 0, 0]


def test_sytrfnp_closed_form():
    r = cbkt.rng(8)
    nm, n, V = 9, 10, 8
    A = cbkt.gen_sym_ldlt(r, nm, n)
    ap = cbkt.pack(A, V)
    assert cbkt.oracle().call("d", "sytrfnp", "C", "L", n, ap, n, V, nm) == 0
    F = cbkt.unpack(ap, nm, n, n, V)
    for k in range(nm):
        L = np.tril(F[k], -1) + np.eye(n)
        D = np.diag(np.diag(F[k]))
        assert np.allclose(L @ D @ L.T, A[k], atol=1e-12 * np.abs(A[k]).max())


def _spd_cond(r, nm, n, kmax):
    G = r.standard_normal((nm, n, n))
    Q, Rr = np.linalg.qr(G)
    lam = 10.0 ** (r.uniform(0, 1, (nm, n)) * kmax)
    A = (Q * lam[:, None, :]) @ np.swapaxes(Q, 1, 2)
    return 0.5 * (A + np.swapaxes(A, 1, 2))


def test_dsposv_oracle_against_lapack():
    r = cbkt.rng(9)
    nm, n, V, nrhs = 18, 16, 8, 2
    A = _spd_cond(r, nm, n, 4)
    A[[3, 7]] = _spd_cond(r, 2, n, 12)     # cond ~ 1e12: FP32 refinement cannot converge
    B = r.standard_normal((nm, n, nrhs))
    ap, bp = cbkt.pack(A, V), cbkt.pack(B, V, pad_diag=0.0)
    xp = np.zeros_like(bp)
    it = np.zeros(nm, np.int32)
    info = np.zeros(nm, np.int32)
    assert cbkt.oracle().call_raw("dsposv_compact", "C", "L", n, nrhs, ap, n, bp, n, xp, n, None, 0,
                                  it, info, V, nm) == 0
    X = cbkt.unpack(xp, nm, n, nrhs, V)
    F = cbkt.unpack(ap, nm, n, n, V)
    for k in range(nm):
        c, x, inf = lapack.dposv(A[k], B[k], lower=1)
        assert inf == 0 and info[k] == 0
        if k in (3, 7):
            assert it[k] < 0
            assert np.allclose(np.tril(F[k]), np.tril(c), atol=1e-10 * np.abs(c).max())
        else:
            assert it[k] >= 0
            assert np.array_equal(F[k], np.where(np.tril(np.ones((n, n))) > 0, A[k], F[k]))
        eta = np.abs(B[k] - A[k] @ X[k]).max() / (np.abs(A[k]).sum(1).max() * np.abs(X[k]).max())
        assert eta <= 1e-15


def _tall_cond(r, nm, m, n, logk):
    """m x n matrices U diag(s) W^T with log10 of the singular values uniform on [0, logk]"""
    U = np.linalg.qr(r.standard_normal((nm, m, n)))[0]
    W = np.linalg.qr(r.standard_normal((nm, n, n)))[0]
    s = 10.0 ** (-r.uniform(0, 1, (nm, n)) * logk)
    s[:, 0] = 1.0
    s[:, -1] = 10.0 ** -logk
    return (U * s[:, None, :]) @ np.swapaxes(W, 1, 2)


def test_dsgels_oracle_against_lapack():
    """the mixed least squares: X against LAPACK dgels; refinement for kappa ~ 10,
    the FP64 path (iter -31, A = its QR) for kappa ~ 1e6, iter -1 for a shape
    without a single-precision kernel; info through the FP64 path"""
    r = cbkt.rng(12)
    for (m, n, nrhs, V) in ((64, 32, 2, 8), (16, 8, 3, 4), (24, 10, 2, 4), (40, 36, 2, 4)):
        nm = 10
        A = r.uniform(-1, 1, (nm, m, n))
        A[[2, 7]] = _tall_cond(r, 2, m, n, 6)
        B = r.standard_normal((nm, m, nrhs))
        ap, bp = cbkt.pack(A, V), cbkt.pack(B, V, pad_diag=0.0)
        bp0 = bp.copy()
        xp = np.zeros(((nm + V - 1) // V) * n * nrhs * V)
        it = np.zeros(nm, np.int32)
        info = np.full(nm, -9, np.int32)
        assert cbkt.oracle().call_raw("dsgels_compact", "C", "N", m, n, nrhs, ap, m, bp, m, xp, n,
                                      it, info, V, nm) == 0
        assert np.array_equal(bp, bp0)
        X = cbkt.unpack(xp, nm, n, nrhs, V)
        F = cbkt.unpack(ap, nm, m, n, V)
        mixed = n <= 32
        for k in range(nm):
            _, x, inf = lapack.dgels(A[k], B[k])[:3]
            assert inf == 0 and info[k] == 0
            s = np.linalg.svd(A[k], compute_uv=False)
            kap = s[0] / s[-1]
            tol = 1e-14 * kap * (1 + kap * np.linalg.norm(B[k] - A[k] @ x[:n]) /
                                 (s[0] * np.linalg.norm(x[:n]))) * np.abs(x[:n]).max() * 50
            assert np.abs(X[k] - x[:n]).max() <= tol, (m, n, k)
            if not mixed:
                assert it[k] == -1
            elif k in (2, 7):
                assert it[k] == -31
                qr = lapack.dgeqrf(A[k])[0]
                assert np.allclose(np.abs(np.triu(F[k])), np.abs(np.triu(qr[:m])), atol=1e-12)
            else:
                assert 0 <= it[k] <= 10
                assert np.array_equal(F[k], A[k])
    # a zero column: FP32 R(j,j) = 0 -> FP64 path -> info j + 1
    m, n, nrhs, V, nm = 32, 16, 1, 4, 4
    A = r.uniform(-1, 1, (nm, m, n))
    A[1, :, 5] = 0.0
    ap, bp = cbkt.pack(A, V), cbkt.pack(r.standard_normal((nm, m, nrhs)), V, pad_diag=0.0)
    xp = np.zeros(n * nrhs * V)
    it = np.zeros(nm, np.int32)
    info = np.zeros(nm, np.int32)
    assert cbkt.oracle().call_raw("dsgels_compact", "C", "N", m, n, nrhs, ap, m, bp, m, xp, n,
                                  it, info, V, nm) == 0
    assert info.tolist() == [0, 6, 0, 0] and it[1] == -31 and np.all(it[[0, 2, 3]] >= 0)
