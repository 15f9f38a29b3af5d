"""GPU parity of the 8 x 8 lane-grid ormqr and orgqr kernels (reflector
lengths, or m for orgqr, above what the register and team kernels serve, up to
64; both precisions, both layouts, both sides and transposes, k from 0 to nq)
against the oracle. A reflector's dot product is summed per lane in row order
and then over the row lanes, where the reference runs one ascending chain, so
the comparison is to rounding: each entry within 20 nq eps max|C| of the
oracle's (orgqr: 20 m eps, Q's entries being at most 1), the leading-dimension
padding untouched, and for orgqr the columns' orthonormality."""

import numpy as np
import pytest

import cases
import cbkt

pytestmark = pytest.mark.gpu


def _gpu():
    import gpuutil
    return gpuutil


RUNS = [("C", 64, 130, 1), ("R", 32, 71, 0), ("C", 8, 13, 1), ("R", 4, 6, 1), ("C", 2, 5, 0)]


@pytest.mark.parametrize("p", ["d", "s"])
@pytest.mark.parametrize("side,m,n,k", [("L", 64, 8, 32), ("L", 48, 5, 24), ("L", 40, 17, 40),
                                        ("L", 28, 3, 28), ("L", 64, 2, 64), ("L", 33, 9, 1),
                                        ("R", 3, 64, 64), ("R", 9, 33, 20), ("R", 1, 48, 0)])
def test_ormqr_grid(p, side, m, n, k):
    for lay, V, nm, lde in RUNS:
        for tr in ("T", "N"):
            print(f"case ormqr{p} {side}{tr} {lay} m={m} n={n} k={k} V={V} nm={nm}", flush=True)
            c = cases.make_case("ormqr", p, {"layout": lay, "side": side, "trans": tr},
                                {"m": m, "n": n, "k": k}, V, nm, seed=9700 + m + 3 * n + k, ld_extra=lde)
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
            if k == 0:
                assert cases.same_bits(co, cg)
            assert np.array_equal(cg == cbkt.SENTINEL, co == cbkt.SENTINEL)


@pytest.mark.parametrize("p", ["d", "s"])
@pytest.mark.parametrize("m,n,k", [(64, 32, 32), (64, 64, 64), (48, 40, 30), (33, 33, 0),
                                   (40, 20, 20), (20, 20, 17), (64, 9, 5), (36, 36, 35)])
def test_orgqr_grid(p, m, n, k):
    for lay, V, nm, lde in RUNS:
        print(f"case orgqr{p} {lay} m={m} n={n} k={k} V={V} nm={nm}", flush=True)
        c = cases.make_case("orgqr", p, {"layout": lay}, {"m": m, "n": n, "k": k}, V, nm,
                            seed=9800 + m + 3 * n + k, ld_extra=lde)
        so, (qo,) = cases.run(cbkt.oracle().call, c)
        sg, (qg,) = cases.run(_gpu().gpu_call, c)
        assert so == sg == 0
        row = lay == "R"
        lda = c["args"][5]
        Qo = cbkt.unpack(qo, nm, m, n, V, ld=lda, row=row).astype(np.float64)
        Qg = cbkt.unpack(qg, nm, m, n, V, ld=lda, row=row).astype(np.float64)
        eps = float(np.finfo(cbkt.DT[p]).eps)
        assert np.abs(Qg - Qo).max() <= 20 * m * eps
        # orthonormal columns
        G = np.swapaxes(Qg, 1, 2) @ Qg
        assert np.abs(G - np.eye(n)).max() <= 20 * m * eps
        if k == 0:
            assert cases.same_bits(qo, qg)
        assert np.array_equal(qg == cbkt.SENTINEL, qo == cbkt.SENTINEL)
