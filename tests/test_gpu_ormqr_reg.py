"""GPU parity of the register ormqr kernel (ormqr_reg.hip: one lane per matrix,
C's columns in registers, reflector length nq <= 24, and <= 32 in FP32, exact
buckets and guarded lengths inside them)
against the oracle, bit for bit: it performs the reference's larf operations
in the lane kernel's order. Both sides, both transposes, both layouts, partial
column blocks (odd ncols), any k <= nq, V from 2 to 64 with a partial last
group; the slots of C outside the matrices stay untouched."""

import numpy as np
import pytest

import cases
import cbkt

pytestmark = pytest.mark.gpu


def _gpu():
    import gpuutil
    return gpuutil


def _bits(x):
    return np.ascontiguousarray(x).view(np.uint8)


@pytest.mark.parametrize("p", ["d", "s"])
@pytest.mark.parametrize("side,m,n,k", [("L", 16, 2, 8), ("L", 8, 5, 8), ("L", 24, 3, 11),
                                        ("R", 3, 16, 16), ("R", 7, 24, 5), ("L", 16, 1, 1),
                                        ("R", 2, 8, 3), ("L", 12, 3, 7), ("R", 4, 20, 9),
                                        ("L", 5, 2, 5), ("L", 1, 3, 1), ("R", 6, 13, 13)])
def test_ormqr_reg(p, side, m, n, k):
    for lay in ("C", "R"):
        for tr in ("T", "N"):
            for V, nm in ((64, 130), (32, 71), (8, 13), (2, 5)):
                print(f"case ormqr{p} {side}{tr} {lay} m={m} n={n} k={k} V={V} nm={nm}", flush=True)
                c = cases.make_case("ormqr", p, {"layout": lay, "side": side, "trans": tr},
                                    {"m": m, "n": n, "k": k}, V, nm, seed=7700 + m + n + k, ld_extra=1)
                so, (co,) = cases.run(cbkt.oracle().call, c)
                sg, (cg,) = cases.run(_gpu().gpu_call, c)
                assert so == sg == 0
                assert np.array_equal(_bits(cg), _bits(co)), (lay, tr, V)


def test_ormqr_reg_fp32_nq32():
    """FP32 at nq = 32 takes this kernel too (FP64 there takes the team kernel)"""
    V, nm = 16, 40
    for side, m, n in (("L", 32, 3), ("R", 5, 32), ("L", 30, 2), ("R", 3, 27)):
        c = cases.make_case("ormqr", "s", {"layout": "C", "side": side, "trans": "T"},
                            {"m": m, "n": n, "k": 20}, V, nm, seed=7800 + m, ld_extra=0)
        so, (co,) = cases.run(cbkt.oracle().call, c)
        sg, (cg,) = cases.run(_gpu().gpu_call, c)
        assert so == sg == 0
        assert np.array_equal(_bits(cg), _bits(co))
