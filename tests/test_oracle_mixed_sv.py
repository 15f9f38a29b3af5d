"""Pin the oracle's mixed-precision symmetric solvers (dsposv_compact and
dssysvnp_compact: single factor, double refinement, double fallback), which
have no reference counterpart, by what the mathematics fixes:

  - X solves A X = B to double backward error on well-conditioned matrices,
    with iter >= 0 and A untouched;
  - cond beyond single precision (1e10 - 1e12: the single factor fails, -3,
    or the refinement cannot converge, -31): A holds the double
    factor, bit-identical to the oracle's own dpotrf / dsytrfnp (both pinned
    against the reference), X is the plain double solve's bit for bit and
    matches numpy's;
  - each ITER code from its construction: a B or A entry beyond FLT_MAX gives
    -2 (B tested first), a diagonal that rounds to 0 in single gives -3 (for
    sysvnp also a D that is finite in double but whose reciprocal overflows
    single), a matrix singular in double as well gives info = j of its first
    failing pivot;
  - a refinement that diverges (an indefinite matrix of cond ~ 1e12) stops
    at -2 once its residual passes FLT_MAX: A and B fit single precision, so
    only the residual test can give that code;
  - nrhs = 0 and n = 0 return iter = info = 0.

A planted mistake (a dropped overflow test, a missing non-finite D test, the
wrong fallback factor) fails one of these.
"""

import numpy as np
import pytest

import cbkt

FLT_MAX = float(np.finfo(np.float32).max)


def _call(name, A, B, V, uplo="L", layout="C"):
    nm, n, _ = A.shape
    nrhs = B.shape[2]
    row = layout == "R"
    ap = cbkt.pack(A, V, row=row)
    ldb = nrhs if row else n
    bp = cbkt.pack(B, V, row=row, pad_diag=0.0)
    xp = np.zeros_like(bp)
    it = np.full(nm, 99, np.int32)
    info = np.full(nm, 99, np.int32)
    st = cbkt.oracle().call_raw(name, layout, uplo, n, nrhs, ap, n, bp, ldb, xp, ldb, None, 0,
                                it, info, V, nm)
    assert st == 0
    return ap, xp, it, info


def _plain(name, A, B, V):
    """X of the oracle's plain double posv / sysvnp (both pinned against the
    reference), the fallback's result bit for bit"""
    nm, n, _ = A.shape
    nrhs = B.shape[2]
    ap, bp = cbkt.pack(A, V), cbkt.pack(B, V, pad_diag=0.0)
    routine = "posv" if name == "dsposv_compact" else "sysvnp"
    cbkt.oracle().call("d", routine, "C", "L", n, nrhs, ap, n, bp, n, V, nm)
    return cbkt.unpack(bp, nm, n, nrhs, V)


def _sym_cond(r, nm, n, logk, definite):
    Q = np.linalg.qr(r.standard_normal((nm, n, n)))[0]
    lam = 10.0 ** (r.uniform(0, 1, (nm, n)) * logk)
    lam[:, 0], lam[:, -1] = 1.0, 10.0 ** logk
    if not definite:
        lam = lam * np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    A = (Q * lam[:, None, :]) @ np.swapaxes(Q, 1, 2)
    return 0.5 * (A + np.swapaxes(A, 1, 2))


def _eta(A, X, B):
    R = B - A @ X
    return np.abs(R).max() / (np.abs(A).sum(1).max() * np.abs(X).max())


@pytest.mark.parametrize("name,definite,logk", [("dsposv_compact", True, 12),
                                                ("dssysvnp_compact", False, 10)])
def test_mixed_sv_oracle_solves_and_falls_back(name, definite, logk):
    r = cbkt.rng(31 if definite else 32)
    nm, n, V, nrhs = 20, 12, 8, 3
    A = cbkt.gen_spd(r, nm, n) if definite else cbkt.gen_sym_ldlt(r, nm, n)
    bad = (4, 11)
    A[list(bad)] = _sym_cond(r, len(bad), n, logk, definite)
    B = r.standard_normal((nm, n, nrhs))
    ap, xp, it, info = _call(name, A, B, V)
    X = cbkt.unpack(xp, nm, n, nrhs, V)
    F = cbkt.unpack(ap, nm, n, n, V)
    assert np.all(info == 0)
    # the double factor the fallback must leave, from the oracle's own plain routine
    fp = cbkt.pack(A, V)
    routine = "potrf" if definite else "sytrfnp"
    assert cbkt.oracle().call("d", routine, "C", "L", n, fp, n, V, nm) == 0
    F64 = cbkt.unpack(fp, nm, n, n, V)
    Xp = _plain(name, A, B, V)
    for k in range(nm):
        if k in bad:
            assert it[k] in (-3, -31), (k, it[k])
            assert np.array_equal(np.tril(F[k]), np.tril(F64[k]))
            assert np.array_equal(X[k], Xp[k])
            Xs = np.linalg.solve(A[k], B[k])
            cond = np.linalg.cond(A[k], 1)
            assert np.abs(X[k] - Xs).max() <= 8 * n * 2.0 ** -53 * cond * np.abs(Xs).max()
        else:
            assert 0 <= it[k] <= 5, (k, it[k])
            assert np.array_equal(np.tril(F[k]), np.tril(A[k]))
            assert _eta(A[k], X[k], B[k]) <= 1e-15


@pytest.mark.parametrize("name", ["dsposv_compact", "dssysvnp_compact"])
def test_mixed_sv_oracle_iter_codes(name):
    ldl = name == "dssysvnp_compact"
    r = cbkt.rng(33)
    nm, n, V, nrhs = 8, 6, 8, 2
    A = cbkt.gen_spd(r, nm, n)
    B = r.standard_normal((nm, n, nrhs))
    B[1, 2, 1] = 1e39                   # B overflows single: -2
    A[2, 4, 1] = A[2, 1, 4] = 1e39      # an off-diagonal entry overflows: -2
    A[2, 4, 4] = 1e80                   # (kept SPD in double)
    A[3, 1, 1] = 1e39                   # both, B tested first: still -2
    B[3, 0, 0] = -1e39
    # a diagonal 1e-50 rounds to 0 in single: the single factor fails (-3)
    A[4] = np.diag(np.full(n, 2.0))
    A[4, 0, 0] = 1e-50
    # singular in double too: pivot 3 is exactly zero (LDL^T) / not positive
    A[5] = np.diag([1.0, 2.0, 0.0, 3.0, 4.0, 5.0])
    if ldl:
        # D(1) = 1e-40 is finite in single, its reciprocal is not: the single
        # factor's later pivots are non-finite (-3); double solves it
        A[6] = np.eye(n) + 0.5 * (np.eye(n, k=1) + np.eye(n, k=-1))
        A[6, 0, 0] = 1e-40
    ap, xp, it, info = _call(name, A, B, V)
    X = cbkt.unpack(xp, nm, n, nrhs, V)
    assert it[1] == -2 and it[2] == -2 and it[3] == -2, it
    assert it[4] == -3 and info[4] == 0, (it, info)
    assert it[5] == -3 and info[5] == 3, (it, info)
    ok = [0, 7] + ([] if ldl else [6])
    assert np.all(it[ok] >= 0) and np.all(info[ok] == 0)
    Xp = _plain(name, A, B, V)
    for k in (1, 2, 3, 4) + ((6,) if ldl else ()):
        assert np.array_equal(X[k], Xp[k]), k
    for k in (2, 4):
        Xs = np.linalg.solve(A[k], B[k])
        assert np.allclose(X[k], Xs, rtol=1e-9, atol=1e-9 * np.abs(Xs).max())
    if ldl:
        assert it[6] == -3 and info[6] == 0, (it, info)


def test_mixed_sv_oracle_residual_overflow():
    r = cbkt.rng(5)
    A = _sym_cond(r, 8, 12, 12, False)
    B = r.standard_normal((8, 12, 3))
    assert np.abs(A).max() < FLT_MAX and np.abs(B).max() < FLT_MAX
    _, xp, it, info = _call("dssysvnp_compact", A, B, 8)
    assert it[4] == -2 and info[4] == 0, it
    X = cbkt.unpack(xp, 8, 12, 3, 8)
    assert np.array_equal(X[4], _plain("dssysvnp_compact", A, B, 8)[4])


@pytest.mark.parametrize("name", ["dsposv_compact", "dssysvnp_compact"])
def test_mixed_sv_oracle_quick_returns(name):
    r = cbkt.rng(34)
    A = cbkt.gen_spd(r, 5, 4)
    _, _, it, info = _call(name, A, np.zeros((5, 4, 0)), 4)
    assert np.all(it == 0) and np.all(info == 0)
    # n = 0, ld = 1: nothing is referenced
    buf = np.zeros(8)
    it = np.full(3, 99, np.int32)
    info = np.full(3, 99, np.int32)
    assert cbkt.oracle().call_raw(name, "C", "L", 0, 2, buf, 1, buf, 1, buf, 1, None, 0, it,
                                  info, 2, 3) == 0
    assert np.all(it == 0) and np.all(info == 0)
