"""GPU parity of the register trsm kernel (trsm_reg.hip: side L, m <= 16, and
m <= 24 in FP32, exact and guarded orders, one lane per matrix with op(A)'s
triangle in registers)
against the oracle, bit for bit: every uplo / trans / diag, both layouts,
alpha 1 and not 1, odd and even right-hand side counts (the last column of an
odd count takes the reference router's axpy form where B has unit row stride
and A's unit stride matches op), V from 2 to 64 with partial last groups; the
rest of B's buffer (leading-dimension padding, padded slots) stays untouched."""

import itertools

import pytest

import cases
import cbkt

pytestmark = pytest.mark.gpu


def _gpu():
    import gpuutil
    return gpuutil


@pytest.mark.parametrize("p,m", [("d", 8), ("s", 8), ("d", 16), ("s", 16), ("s", 24), ("s", 1), ("d", 5),
                                 ("s", 12), ("d", 13), ("s", 20)])
def test_trsm_reg(p, m):
    combos = list(itertools.product("CR", "UL", "NT", "UN"))
    for i, (lay, uplo, tr, diag) in enumerate(combos):
        for n, alpha, V, nm in ((3, 1.0, 64, 130), (4, -0.75, 8, 13), (1, 1.0, 2, 5), (m, 0.5, 16, 20), (2, 1.0, 32, 71)):
            print(f"case trsm{p} {lay} L{uplo}{tr}{diag} m={m} n={n} alpha={alpha} V={V} nm={nm}", flush=True)
            c = cases.make_case("trsm", p, {"layout": lay, "side": "L", "uplo": uplo, "trans": tr, "diag": diag},
                                {"m": m, "n": n, "alpha": alpha}, V, nm, seed=9300 + 17 * m + i + n,
                                ld_extra=(i + n) % 2)
            so, (bo,) = cases.run(cbkt.oracle().call, c)
            sg, (bg,) = cases.run(_gpu().gpu_call, c)
            assert so == sg == 0
            assert cases.same_bits(bo, bg), (lay, uplo, tr, diag, n, alpha, V)
