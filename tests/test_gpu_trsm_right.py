"""GPU parity of side-R trsm (X op(A) = alpha B) against the oracle, bit for
bit. Side R runs the side-L register, quad and team kernels on B's
transposed view with trans flipped (op(A)^T X^T = alpha B^T): per element
of X that is the reference's side-R sequence (the multiply by alpha, the
fma's over l ascending, the divide), so the bits must match. Every
layout / uplo / trans / diag at the orders each kernel takes (register:
n <= 16 FP64, <= 24 FP32; quad: 24 FP64, 32 FP32; team: 16 / 24 / 32), odd
and even row counts, V from 4 to 64 with partial last groups; B's padding
stays untouched."""

import itertools

import pytest

import cases
import cbkt

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("p,n", [("d", 8), ("d", 16), ("d", 20), ("d", 24), ("d", 32), ("s", 20),
                                 ("s", 24), ("s", 28), ("s", 32), ("d", 5), ("d", 40), ("s", 64)])
def test_trsm_side_r_matches_oracle(p, n):
    import gpuutil
    combos = list(itertools.product("CR", "UL", "NT", "UN"))
    for i, (lay, uplo, tr, diag) in enumerate(combos):
        for m, alpha, V, nm in ((24, 1.0, 64, 130), (3, -0.75, 8, 13), (n, 0.5, 32, 70), (1, 2.0, 4, 9)):
            print(f"case trsm{p} {lay} R{uplo}{tr}{diag} m={m} n={n} alpha={alpha} V={V} nm={nm}",
                  flush=True)
            c = cases.make_case("trsm", p, {"layout": lay, "side": "R", "uplo": uplo, "trans": tr,
                                            "diag": diag},
                                {"m": m, "n": n, "alpha": alpha}, V, nm, seed=9700 + 13 * n + i + m,
                                ld_extra=(i + m) % 2)
            so, (bo,) = cases.run(cbkt.oracle().call, c)
            sg, (bg,) = cases.run(gpuutil.gpu_call, c)
            assert so == sg == 0
            assert cases.same_bits(bo, bg), (lay, uplo, tr, diag, m, alpha, V)
