"""The oracle's uneven-batch gels (no reference counterpart): pinned to the
oracle's compact gels on each problem packed alone (which the reference and
LAPACK pin), bit for bit, and to LAPACK dgels directly; the quick returns
(nrhs = 0 leaves everything, min(m, n) = 0 zeroes B) and info codes."""

import numpy as np
import scipy.linalg.lapack as lapack

import batchcase
import cases
import cbkt


def _compact_one(p, trans, A, B):
    """gels_compact of the oracle on one matrix (V = 2, slot 1 identity)"""
    m, n = A.shape
    mx = max(m, n)
    nrhs = B.shape[1]
    ap = cbkt.pack(A[None], 2)
    bp = cbkt.pack(B[None], 2, pad_diag=0.0)
    tp = np.zeros(2 * max(1, min(m, n)), A.dtype)
    info = np.zeros(1, np.int32)
    assert cbkt.oracle().call(p, "gels", "C", trans, m, n, nrhs, ap, max(1, m), bp, max(1, mx),
                              tp, 2, 1, info=info) == 0
    return cbkt.unpack(ap, 1, m, n, 2)[0], cbkt.unpack(bp, 1, mx, nrhs, 2)[0], int(info[0])


def test_batch_equals_compact_per_problem():
    r = cbkt.rng(71)
    for p in ("d", "s"):
        for trans in ("N", "T"):
            shapes = batchcase.uneven_shapes(r, 40, mmax=20, rmax=3)
            shapes += [(5, 9, 2), (1, 1, 1), (7, 3, 1)]
            desc, ap, bp, As, Bs = batchcase.make(r, shapes, cbkt.DT[p])
            info = np.full(len(desc), -9, np.int32)
            assert batchcase.oracle_batch(p, trans, desc, ap, bp, info) == 0
            for k, d in enumerate(desc):
                Fc, Xc, ic = _compact_one(p, trans, As[k], Bs[k])
                assert cases.same_bits(batchcase.get_a(ap, d), Fc)
                assert cases.same_bits(batchcase.get_b(bp, d), Xc)
                assert info[k] == ic


def test_batch_against_lapack_dgels():
    r = cbkt.rng(72)
    shapes = [(m, n, q) for m, n, q in batchcase.uneven_shapes(r, 60, mmax=40)]
    desc, ap, bp, As, Bs = batchcase.make(r, shapes)
    info = np.zeros(len(desc), np.int32)
    assert batchcase.oracle_batch("d", "N", desc, ap, bp, info) == 0
    assert np.all(info == 0)
    for k, d in enumerate(desc):
        m, n, q = d[:3]
        _, x, inf = lapack.dgels(As[k], Bs[k])[:3]
        assert inf == 0
        X = batchcase.get_b(bp, d)
        assert np.allclose(X[:n], x[:n], rtol=1e-9, atol=1e-11), (k, d)


def test_batch_quick_returns_and_info():
    r = cbkt.rng(73)
    shapes = [(6, 4, 0), (0, 3, 2), (5, 0, 2), (8, 4, 2), (9, 5, 1)]
    desc, ap, bp, As, Bs = batchcase.make(r, shapes)
    # problem 3: a zero column -> info 3 (first zero R(i,i)), B = Q^T B
    m, n, q, lda, ldb, a0, b0 = desc[3]
    ap[a0 + 2 * lda: a0 + 2 * lda + m] = 0.0
    A3 = batchcase.get_a(ap, desc[3]).copy()
    ap0, bp0 = ap.copy(), bp.copy()
    info = np.full(len(desc), -9, np.int32)
    assert batchcase.oracle_batch("d", "N", desc, ap, bp, info) == 0
    assert info.tolist() == [0, 0, 0, 3, 0]
    # nrhs = 0: nothing touched
    assert np.array_equal(batchcase.get_a(ap, desc[0]), batchcase.get_a(ap0, desc[0]))
    # min(m, n) = 0: B := 0 over max(m, n) rows
    assert np.all(batchcase.get_b(bp, desc[1]) == 0) and np.all(batchcase.get_b(bp, desc[2]) == 0)
    # the failed problem holds Q^T B: its norm equals B's (Q orthogonal)
    B3 = batchcase.get_b(bp0, desc[3])
    assert np.allclose(np.linalg.norm(batchcase.get_b(bp, desc[3]), axis=0),
                       np.linalg.norm(B3, axis=0))
    _, _, _, inf = lapack.dgeqrf(A3)
    # padding between problems untouched
    mask = np.ones(len(ap), bool)
    for d in desc:
        m, n, q, lda, ldb, a0, b0 = d
        for j in range(n):
            mask[a0 + j * lda: a0 + j * lda + m] = False
    assert np.all(ap[mask] == 7.0)
