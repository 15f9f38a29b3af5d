"""GPU parity of the register lane kernel (potrf/posv/potrs, sytrfnp/sysvnp/
sytrsnp at n <= 16 in FP64, n <= 24 in FP32) against the oracle, bit for bit
-- factors and solutions:
the kernel performs the lane kernel's operations in registers, including the
reference trsm router's choice of the axpy form for a single leftover
right-hand side of the backward sweep (nrhs = 1, 3, 5 with column-major B).
Also the per-matrix info codes with a failed matrix keeping its B."""

import numpy as np
import pytest

import cases
import cbkt

pytestmark = pytest.mark.gpu


def _gpu():
    import gpuutil
    return gpuutil


@pytest.mark.parametrize("p,n", [(p, n) for p in "ds" for n in (1, 3, 8, 9, 12, 13, 16)] +
                         [("s", 20), ("s", 24)])
def test_reg_factor_and_solve_bitwise(p, n):
    for i, (lay, uplo) in enumerate((("C", "L"), ("C", "U"), ("R", "L"), ("R", "U"))):
        for V, nm in ((64, 130), (8, 13), (2, 5)):
            for rt in ("potrf", "sytrfnp", "posv", "sysvnp", "potrs", "sytrsnp"):
                for nrhs in ((0,) if rt in ("potrf", "sytrfnp") else (1, 3, 4, 5)):
                    dims = {"n": n} if nrhs == 0 else {"n": n, "nrhs": nrhs}
                    print(f"case reg {rt}{p} {lay}{uplo} n={n} nrhs={nrhs} V={V} nm={nm}", flush=True)
                    c = cases.make_case(rt, p, {"layout": lay, "uplo": uplo}, dims, V, nm,
                                        seed=9100 + 7 * n + i + nrhs, ld_extra=i % 2)
                    so, outo = cases.run(cbkt.oracle().call, c)
                    sg, outg = cases.run(_gpu().gpu_call, c)
                    assert so == sg == 0
                    for a, b in zip(outo, outg):
                        assert cases.same_bits(a, b), (rt, lay, uplo, V, nrhs)


def test_reg_info():
    V, nm, n = 64, 70, 8
    c = cases.make_case("posv", "d", {"layout": "C", "uplo": "L"}, {"n": n, "nrhs": 2}, V, nm, seed=4)
    A = cbkt.unpack(c["args"][4], nm, n, n, V)
    A[7, 5, 5] = -3.0
    c["args"][4] = cbkt.pack(A, V)
    info_o = np.zeros(nm, np.int32)
    info_g = np.zeros(nm, np.int32)
    so, outo = cases.run(cbkt.oracle().call, c, info=info_o)
    sg, outg = cases.run(_gpu().gpu_call, c, info=info_g)
    assert so == sg == 0
    assert np.array_equal(info_o, info_g) and info_o[7] == 6
    for a, b in zip(outo, outg):
        assert cases.same_bits_nan(a, b)
