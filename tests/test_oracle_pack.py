"""Pin the oracle's dense <-> compact conversion (cbk_?pack_compact and
friends, restating the reference's pack_compact / unpack_compact of
tests/test_compact_util.hpp:475-531):

  - against the reference itself, compiled from its own sources
    (oracle/_ref/libcbkref_pack.so, column-major dense, V in {2,4,8,16}),
    bit for bit, padded slots included (skipped where that build is absent);
  - against cbkt.pack, an independent numpy restatement, for every V, both
    layouts, a dense side in the other layout, padded leading dimensions and
    batch strides: the compact buffer bit for bit, the leading-dimension
    padding untouched, the padded slots the identity (or pad_diag);
  - unpack inverts pack on the real slots and leaves the dense padding alone;
  - the pointer-array forms equal the strided ones;
  - argument codes in argument order.
"""

import ctypes
import os

import numpy as np
import pytest

import cbkt
import packcase


def run_pack(lib, M, layout, alayout, lda, ldap, stride, V, pad=1.0, fn="pack"):
    nm, rows, cols = M.shape
    p = "d" if M.dtype == np.float64 else "s"
    a = packcase.dense_batch(M, alayout, lda, stride)
    ap = np.full(cbkt.buf_shape(nm, rows, cols, ldap, V, layout == "R"), cbkt.SENTINEL,
                 dtype=M.dtype).reshape(-1)
    st = lib.call_raw(f"cbk_{p}pack_compact", layout, rows, cols, a, alayout, lda, stride, ap, ldap,
                      pad, V, nm)
    return st, ap


@pytest.mark.parametrize("case", packcase.CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_pack_matches_numpy_restatement(case, dtype):
    layout, alayout, rows, cols, *_ , V, nm = case
    M, lda, ldap, stride = packcase.make(case, dtype, 11)
    o = cbkt.oracle()
    for pad in (1.0, 0.0):
        st, ap = run_pack(o, M, layout, alayout, lda, ldap, stride, V, pad)
        assert st == 0
        want = cbkt.pack(M, V, ld=ldap, row=layout == "R", pad_diag=pad, dtype=dtype)
        assert np.array_equal(ap, want)
    # unpack inverts it; the dense buffer's padding is left alone
    p = "d" if dtype == np.float64 else "s"
    a = np.full_like(packcase.dense_batch(M, alayout, lda, stride), cbkt.SENTINEL)
    assert o.call_raw(f"cbk_{p}unpack_compact", layout, rows, cols, ap, ldap, a, alayout, lda,
                      stride, V, nm) == 0
    assert np.array_equal(packcase.from_dense(a, nm, rows, cols, alayout, lda, stride), M)
    assert np.array_equal(a, packcase.dense_batch(M, alayout, lda, stride))


@pytest.mark.parametrize("case", [c for c in packcase.CASES if c[0] == c[1]])
def test_pointer_array_forms_equal_strided(case):
    layout, alayout, rows, cols, *_ , V, nm = case
    M, lda, ldap, stride = packcase.make(case, np.float64, 12)
    o = cbkt.oracle()
    a = packcase.dense_batch(M, alayout, lda, stride)
    ptrs = np.array([a.ctypes.data + 8 * k * stride for k in range(nm)], dtype=np.uint64)
    ap = np.full(cbkt.buf_shape(nm, rows, cols, ldap, V, layout == "R"), cbkt.SENTINEL).reshape(-1)
    assert o.call_raw("cbk_dgepack_compact", layout, rows, cols, ptrs, lda, ap, ldap, V, nm) == 0
    assert np.array_equal(ap, run_pack(o, M, layout, alayout, lda, ldap, stride, V)[1])
    w = np.full_like(a, cbkt.SENTINEL)
    wptrs = np.array([w.ctypes.data + 8 * k * stride for k in range(nm)], dtype=np.uint64)
    assert o.call_raw("cbk_dgeunpack_compact", layout, rows, cols, wptrs, lda, ap, ldap, V, nm) == 0
    assert np.array_equal(w, a)


def _ref_pack_lib():
    path = os.path.join(cbkt.REF_DIR, "libcbkref_pack.so")
    if not os.path.exists(path):
        pytest.skip("the reference's pack utilities are not built (make -C oracle ref)")
    return ctypes.CDLL(path)


@pytest.mark.parametrize("V", [2, 4, 8, 16])
@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_pack_matches_reference_build(V, layout, dtype):
    lib = _ref_pack_lib()
    r = cbkt.rng(13 + V)
    for rows, cols, nm, extra in ((5, 3, 2 * V + 1, 0), (8, 8, V, 2), (1, 7, 3, 1)):
        M = r.standard_normal((nm, rows, cols)).astype(dtype)
        ldap = (cols if layout == "R" else rows) + extra
        p = "d" if dtype == np.float64 else "s"
        ct = ctypes.c_double if p == "d" else ctypes.c_float
        # the reference packs from a column-major batch with ld = rows into a zeroed buffer
        dense = np.ascontiguousarray(np.swapaxes(M, 1, 2)).reshape(-1)
        shape = cbkt.buf_shape(nm, rows, cols, ldap, V, layout == "R")
        want = np.zeros(shape, dtype=dtype).reshape(-1)
        fn = getattr(lib, f"refpack_{p}")
        fn.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 2 + [ctypes.c_int, ct] + [ctypes.c_int] * 2
        for pad in (1.0, 0.0):
            want[:] = 0
            assert fn(layout == "R", rows, cols, dense.ctypes.data, want.ctypes.data, ldap, pad, V, nm) == 0
            got = np.zeros_like(want)
            assert cbkt.oracle().call_raw(f"cbk_{p}pack_compact", layout, rows, cols, dense, "C",
                                          rows, rows * cols, got, ldap, pad, V, nm) == 0
            assert np.array_equal(got, want)
        if p == "d":
            un = lib.refunpack_d
            un.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p] + [ctypes.c_int] * 2
            back = np.zeros_like(dense)
            assert un(layout == "R", rows, cols, want.ctypes.data, ldap, back.ctypes.data, V, nm) == 0
            mine = np.zeros_like(dense)
            assert cbkt.oracle().call_raw("cbk_dunpack_compact", layout, rows, cols, want, ldap, mine,
                                          "C", rows, rows * cols, V, nm) == 0
            assert np.array_equal(mine, back) and np.array_equal(back, dense)


def test_pack_argument_codes():
    o = cbkt.oracle()
    z = np.zeros(64)
    ok = dict(layout="C", rows=4, cols=3, a=z, alayout="C", lda=4, stride=12, ap=z, ldap=4,
              pad=1.0, V=4, nm=2)
    def code(**kw):
        d = dict(ok, **kw)
        return o.call_raw("cbk_dpack_compact", d["layout"], d["rows"], d["cols"], d["a"], d["alayout"],
                          d["lda"], d["stride"], d["ap"], d["ldap"], d["pad"], d["V"], d["nm"])
    assert code() == 0
    assert code(layout="X") == -1
    assert code(rows=-1) == -2
    assert code(cols=-1) == -3
    assert code(a=None) == -4
    assert code(alayout="Q") == -5
    assert code(lda=3) == -6
    assert code(stride=11) == -7
    assert code(ap=None) == -8
    assert code(ldap=3) == -9
    assert code(V=3) == -11
    assert code(nm=-1) == -12
    assert code(a=None, ap=None, nm=0) == 0      # nothing to read or write
    assert code(stride=0, nm=1) == 0             # the stride is unused for one matrix
