"""GPU parity of the ormqr team kernel (reflector length nq in {16, 32}, both
sides, both transposes, any k <= nq) against the oracle. The reflector sums are
butterflies over the team where the reference sums rows in ascending order, so
the comparison is to rounding: op(Q) is orthogonal, and each entry of the
result is held to |gpu - oracle| <= 20 nq eps max|C| per matrix."""

import numpy as np
import pytest

import cases
import cbkt

pytestmark = pytest.mark.gpu


def _gpu():
    import gpuutil
    return gpuutil


@pytest.mark.parametrize("p", ["d", "s"])
@pytest.mark.parametrize("side,m,n,k", [("L", 16, 2, 8), ("L", 32, 5, 16), ("L", 16, 9, 16),
                                        ("R", 3, 16, 8), ("R", 7, 32, 32), ("L", 32, 1, 1)])
def test_ormqr_team(p, side, m, n, k):
    for lay in ("C", "R"):
        for tr in ("T", "N"):
            for V, nm in ((64, 130), (32, 71), (8, 13), (4, 6)):
                print(f"case ormqr{p} {side}{tr} {lay} m={m} n={n} k={k} V={V} nm={nm}", flush=True)
                c = cases.make_case("ormqr", p, {"layout": lay, "side": side, "trans": tr},
                                    {"m": m, "n": n, "k": k}, V, nm, seed=9900 + m + n + k, ld_extra=1)
                so, (co,) = cases.run(cbkt.oracle().call, c)
                sg, (cg,) = cases.run(_gpu().gpu_call, c)
                assert so == sg == 0
                row = lay == "R"
                ldc = c["args"][10]
                Co = cbkt.unpack(co, nm, m, n, V, ld=ldc, row=row).astype(np.float64)
                Cg = cbkt.unpack(cg, nm, m, n, V, ld=ldc, row=row).astype(np.float64)
                C0 = cbkt.unpack(c["args"][9], nm, m, n, V, ld=ldc, row=row).astype(np.float64)
                eps = float(np.finfo(cbkt.DT[p]).eps)
                nq = m if side == "L" else n
                tol = 20 * nq * eps * np.abs(C0).max(axis=(1, 2))
                assert np.all(np.abs(Cg - Co).max(axis=(1, 2)) <= tol)
                # the leading-dimension padding is untouched
                assert np.array_equal(cg == cbkt.SENTINEL, co == cbkt.SENTINEL)
