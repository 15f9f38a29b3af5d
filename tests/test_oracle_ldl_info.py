"""The oracle's unpivoted LDL^T info codes (sytrfnp_compact_info,
sysvnp_compact_info) against their closed form.

LAPACK has no unpivoted LDL^T, so the code is pinned by construction: a
symmetric matrix with a known-good factorization (|D| >= 0.5) whose row and
column k are zeroed up to the diagonal, A(k, 0:k+1) = A(0:k+1, k) = 0, has
L(k, j) = 0 for j < k and hence D(k) = A(k, k) - sum L(k, j)^2 D(j) = 0
exactly, with D(0..k-1) untouched: the first zero pivot is k and info = k + 1
(the reference's convention, src/cbk_sytrfnp_compact.hpp:216-235 parks
1 / D(k) = inf at that pivot). Every k of an order, both uplos, both layouts,
both precisions; sysvnp leaves that matrix's B unmodified; every other
matrix reports 0. Cholesky's codes are pinned against LAPACK dpotrf in
test_oracle_lapack.py.
"""

import numpy as np
import pytest

import cbkt


@pytest.mark.parametrize("p", ["d", "s"])
@pytest.mark.parametrize("layout,uplo", [("C", "L"), ("C", "U"), ("R", "L"), ("R", "U")])
@pytest.mark.parametrize("n", [1, 5, 8, 9, 24])
def test_ldl_info_is_first_zero_pivot(p, layout, uplo, n):
    dt = cbkt.DT[p]
    V = 8
    nm = n + 3  # matrix k has its zero minor at k; the last 3 are regular
    r = cbkt.rng(61000 + 10 * n + (layout == "R") * 2 + (uplo == "U"))
    A = cbkt.gen_sym_ldlt(r, nm, n, np.float64)
    for k in range(n):
        A[k, k, :k + 1] = 0.0
        A[k, :k + 1, k] = 0.0
    A = A.astype(dt)
    row = layout == "R"
    ap = cbkt.pack(A, V, row=row, dtype=dt)
    info = np.full(nm, -7, np.int32)
    assert cbkt.oracle().call(p, "sytrfnp", layout, uplo, n, ap, n, V, nm, info=info) == 0
    want = np.array([k + 1 for k in range(n)] + [0, 0, 0], np.int32)
    assert np.array_equal(info, want), (info, want)

    nrhs = 3
    ap = cbkt.pack(A, V, row=row, dtype=dt)
    B = r.standard_normal((nm, n, nrhs)).astype(dt)
    ldb = nrhs if row else n
    bp = cbkt.pack(B, V, ld=ldb, row=row, pad_diag=0.0, dtype=dt)
    b0 = bp.copy()
    info = np.full(nm, -7, np.int32)
    assert cbkt.oracle().call(p, "sysvnp", layout, uplo, n, nrhs, ap, n, bp, ldb, V, nm,
                              info=info) == 0
    assert np.array_equal(info, want)
    Bg = cbkt.unpack(bp, nm, n, nrhs, V, ld=ldb, row=row)
    B0 = cbkt.unpack(b0, nm, n, nrhs, V, ld=ldb, row=row)
    assert np.array_equal(Bg[:n], B0[:n])  # failed matrices keep their B
    # the regular ones are solved: A x = b to the precision's rounding
    for m in range(n, nm):
        res = A[m].astype(np.float64) @ Bg[m].astype(np.float64) - B0[m].astype(np.float64)
        an = np.abs(A[m].astype(np.float64)).sum(axis=1).max()
        assert np.abs(res).max() <= 100 * n * np.finfo(dt).eps * an * np.abs(Bg[m]).max()
