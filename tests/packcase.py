"""Shared cases of the dense <-> compact conversion tests: a dense batch in
a given layout, leading dimension and batch stride, and the compact buffer
cbkt.pack builds from it (the numpy restatement of the reference's
pack_compact, tests/test_compact_util.hpp:475-506)."""

import numpy as np

import cbkt


def dense_batch(M, alayout, lda, stride):
    """M[nm, rows, cols] laid out as a strided batch (flat array), the padding
    of each leading dimension and of the stride holding a sentinel"""
    nm, rows, cols = M.shape
    buf = np.full(max(1, (nm - 1) * stride + lda * (rows if alayout == "R" else cols)),
                  cbkt.SENTINEL, dtype=M.dtype)
    for k in range(nm):
        blk = M[k] if alayout == "R" else M[k].T
        view = buf[k * stride:k * stride + blk.shape[0] * lda].reshape(blk.shape[0], lda)
        view[:, :blk.shape[1]] = blk
    return buf


def from_dense(buf, nm, rows, cols, alayout, lda, stride):
    out = np.empty((nm, rows, cols), dtype=buf.dtype)
    for k in range(nm):
        n0 = rows if alayout == "R" else cols
        view = buf[k * stride:k * stride + n0 * lda].reshape(n0, lda)
        blk = view[:, :(cols if alayout == "R" else rows)]
        out[k] = blk if alayout == "R" else blk.T
    return out


# (layout, alayout, rows, cols, lda_pad, ldap_pad, stride_pad, V, nm)
CASES = [
    ("C", "C", 5, 3, 0, 0, 0, 4, 9),
    ("C", "C", 32, 32, 0, 0, 0, 64, 130),
    ("R", "R", 7, 4, 2, 1, 5, 8, 17),
    ("C", "R", 6, 9, 1, 3, 0, 2, 5),
    ("R", "C", 16, 8, 0, 2, 3, 16, 33),
    ("C", "C", 1, 1, 0, 0, 0, 32, 31),
    ("R", "C", 24, 24, 3, 0, 0, 32, 65),
    ("C", "C", 64, 2, 0, 0, 1, 64, 64),
    ("C", "C", 3, 70, 0, 5, 0, 2, 3),
]


def make(case, dtype, seed):
    layout, alayout, rows, cols, lp, lpp, sp, V, nm = case
    r = cbkt.rng(seed)
    M = r.standard_normal((nm, rows, cols)).astype(dtype)
    lda = (cols if alayout == "R" else rows) + lp
    ldap = (cols if layout == "R" else rows) + lpp
    stride = lda * (rows if alayout == "R" else cols) + sp
    return M, lda, ldap, stride
