"""GPU parity of the default trsm kernels at m in {16, 20, 24, 28, 32, 40, 64}
(the team and quad kernels on side L, forward sweeps: uplo L without
transpose, uplo U transposed; FP64 20 and FP32 28 on the quad kernel's
instances of order 24 / 32 as diag(op(A), I); 40 and 64 on the 8 x 8 lane
grid) against the oracle, bit for bit: the kernel performs the reference's operations on every element in its
order (alpha first, the fma's over l ascending, then the divide). The backward
sweeps and side R, which the lane kernel serves, are checked alongside."""

import pytest

import cases
import cbkt

pytestmark = pytest.mark.gpu


def _gpu():
    import gpuutil
    return gpuutil


@pytest.mark.parametrize("p", ["d", "s"])
@pytest.mark.parametrize("m", [16, 20, 24, 28, 32, 40, 64])
def test_trsm_team_bitwise(p, m):
    for lay in ("C", "R"):
        for uplo, tr in (("L", "N"), ("U", "T"), ("U", "C"), ("L", "T"), ("U", "N")):
            for diag in ("U", "N"):
                for n, alpha, V, nm in ((24, 1.0, 64, 130), (1, -0.5, 8, 13), (5, 2.0, 4, 6), (4, 1.0, 64, 64), (3, -1.0, 32, 71)):
                    print(f"case trsm{p} L{uplo}{tr}{diag} {lay} m={m} n={n} alpha={alpha} V={V} nm={nm}",
                          flush=True)
                    c = cases.make_case("trsm", p, {"layout": lay, "side": "L", "uplo": uplo,
                                                    "trans": tr, "diag": diag},
                                        {"m": m, "n": n, "alpha": alpha}, V, nm,
                                        seed=8000 + m + n + V, ld_extra=(n % 2))
                    so, outo = cases.run(cbkt.oracle().call, c)
                    sg, outg = cases.run(_gpu().gpu_call, c)
                    assert so == sg == 0
                    for a, b in zip(outo, outg):
                        assert cases.same_bits(a, b), (lay, uplo, tr, diag, n, V)
