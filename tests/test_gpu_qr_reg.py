"""GPU parity of the register QR kernels (qr_reg.hip: geqrf, the
overdetermined gels and orgqr with one lane per matrix and the tall view in registers,
p <= 16 with q <= 8, p <= 32 with q <= 4 (q <= 8 in FP32), exact register
buckets and guarded shapes inside them) against the oracle,
bit for bit: it performs the lane kernel's operations in the same order, so
every output buffer (factor, tau, solution and residual rows, the leading-
dimension padding and the padded group slots) must equal the oracle's exactly.
Both layouts, transposed (wide A, trans T) and plain problems, odd and even
right-hand side counts (the last column of an odd count takes trsm's axpy form
where both views have unit row stride), V from 2 to 64 with partial last
groups, and the info path."""

import numpy as np
import pytest

import cases
import cbkt

pytestmark = pytest.mark.gpu


def _gpu():
    import gpuutil
    return gpuutil


def _check(case):
    so, outo = cases.run(cbkt.oracle().call, case)
    sg, outg = cases.run(_gpu().gpu_call, case)
    assert so == sg == 0
    for a, b in zip(outo, outg):
        assert cases.same_bits(a, b), (case["routine"], case["p"], case["opts"], case["dims"],
                                       case["V"], case["nm"])


# exact buckets, then guarded shapes inside them
SHAPES = [("d", 16, 8), ("s", 16, 8), ("d", 8, 4), ("s", 8, 8), ("d", 8, 8), ("d", 16, 4),
          ("s", 16, 4), ("d", 32, 4), ("s", 32, 4), ("s", 32, 8), ("s", 8, 4),
          ("d", 12, 5), ("s", 7, 3), ("d", 30, 3), ("s", 25, 7), ("d", 16, 7), ("d", 5, 5),
          ("s", 13, 8), ("d", 3, 2)]


@pytest.mark.parametrize("p,m,n", SHAPES)
def test_qr_reg_gels(p, m, n):
    for i, (layout, trans, nrhs, V, nm) in enumerate([("C", "N", 2, 64, 130), ("C", "N", 3, 8, 13),
                                                      ("R", "N", 1, 4, 9), ("C", "T", 5, 2, 5),
                                                      ("R", "T", 2, 64, 70), ("C", "N", 1, 2, 3),
                                                      ("R", "T", 1, 8, 11), ("C", "N", 2, 32, 71)]):
        am, an = (m, n) if trans == "N" else (n, m)
        print(f"case gels{p} {am}x{an} nrhs={nrhs} {layout}{trans} V={V} nm={nm}", flush=True)
        _check(cases.make_case("gels", p, {"layout": layout, "trans": trans},
                               {"m": am, "n": an, "nrhs": nrhs}, V, nm, seed=8100 + 31 * m + n + i,
                               ld_extra=i % 2))


@pytest.mark.parametrize("p,m,n", SHAPES)
def test_qr_reg_geqrf(p, m, n):
    for i, (layout, V, nm) in enumerate([("C", 64, 130), ("R", 8, 11), ("C", 2, 3), ("R", 32, 71)]):
        print(f"case geqrf{p} {m}x{n} {layout} V={V} nm={nm}", flush=True)
        _check(cases.make_case("geqrf", p, {"layout": layout}, {"m": m, "n": n}, V, nm,
                               seed=8300 + 31 * m + n + i, ld_extra=1 - i % 2))


def test_qr_reg_gels_info():
    """exact zero columns: info = their index + 1, B left as Q^T B, the other
    matrices solved normally; bit for bit with the oracle (LAPACK dgels)"""
    V, nm, m, n, nrhs = 16, 40, 16, 8, 3
    c = cases.make_case("gels", "d", {"layout": "C", "trans": "N"}, {"m": m, "n": n, "nrhs": nrhs},
                        V, nm, seed=61)
    A = cbkt.unpack(c["args"][5], nm, m, n, V)
    A[3, :, 7] = 0.0
    A[20, :, 0] = 0.0
    c["args"][5] = cbkt.pack(A, V)
    info_o = np.zeros(nm, np.int32)
    info_g = np.zeros(nm, np.int32)
    so, outo = cases.run(cbkt.oracle().call, c, info=info_o)
    sg, outg = cases.run(_gpu().gpu_call, c, info=info_g)
    assert so == sg == 0
    assert info_o[3] == 8 and info_o[20] == 1 and np.count_nonzero(info_o) == 2
    assert np.array_equal(info_o, info_g)
    for a, b in zip(outo, outg):
        assert cases.same_bits(a, b)


@pytest.mark.parametrize("p,m,n", SHAPES)
def test_qr_reg_orgqr(p, m, n):
    """orgqr of the same shapes: all k from 0 to n over the cases, bit for bit"""
    for i, (layout, k, V, nm) in enumerate([("C", n, 64, 130), ("R", n - 1, 8, 11), ("C", n // 2, 2, 3),
                                            ("R", 0, 4, 9), ("C", 1, 16, 17), ("C", n, 32, 71)]):
        print(f"case orgqr{p} {m}x{n} k={k} {layout} V={V} nm={nm}", flush=True)
        _check(cases.make_case("orgqr", p, {"layout": layout}, {"m": m, "n": n, "k": k}, V, nm,
                               seed=8500 + 31 * m + n + i, ld_extra=i % 2))
