"""GPU parity of the default kernels (potrf/posv/potrs, sytrfnp/sysvnp/sytrsnp)
at orders 16 to 64 against the oracle: the quad kernels at 24 and 32, and at
20 and 28 their instances of order 24 and 32 running the matrix as
diag(A, I) (FP64; FP32 order 20 is the register kernel's); the 8 x 8 lane
grid (one matrix per wavefront) at 33, 40 and 64.

Factors are bit-identical to the oracle's (hence the reference's). Solutions
share that factor; their backward substitution sums in a different order, so
X is held to the solution tolerance of the spec: per right-hand side,
    ||x_gpu - x_oracle||_inf / ||x_oracle||_inf <= 4 * n * eps * cond_1(A)
with cond_1(A) = ||A||_1 ||A^-1||_1 of the dense input.
Also checked: the opposite triangle and the leading-dimension padding stay
untouched, posv == potrf + potrs bit for bit within the library, and the
per-matrix info codes.
"""

import numpy as np
import pytest

import cases
import cbkt

pytestmark = pytest.mark.gpu

TEAM_N = (16, 20, 24, 28, 32, 33, 40, 64)


def _gpu():
    import gpuutil
    return gpuutil


def _cond1(A):
    return np.linalg.norm(A, 1, axis=(1, 2)) * np.linalg.norm(np.linalg.inv(A), 1, axis=(1, 2))


def _dense_sym(case):
    """the symmetric input matrices of a posv/sysvnp case, from the referenced triangle"""
    layout, uplo = case["opts"]["layout"], case["opts"]["uplo"]
    n, V, nm = case["dims"]["n"], case["V"], case["nm"]
    M = cbkt.unpack(case["args"][4], nm, n, n, V, ld=case["args"][5],
                    row=layout == "R").astype(np.float64)
    T = np.tril(M) if uplo == "L" else np.triu(M)
    return T + np.swapaxes(T, 1, 2) - np.einsum("bii->bi", T)[:, :, None] * np.eye(n)


def _check_solution(case, xo, xg):
    p = case["p"]
    n, nrhs, V, nm = case["dims"]["n"], case["dims"]["nrhs"], case["V"], case["nm"]
    layout = case["opts"]["layout"]
    ldb = case["args"][7]
    Xo = cbkt.unpack(xo, nm, n, nrhs, V, ld=ldb, row=layout == "R").astype(np.float64)
    Xg = cbkt.unpack(xg, nm, n, nrhs, V, ld=ldb, row=layout == "R").astype(np.float64)
    eps = float(np.finfo(cbkt.DT[p]).eps)
    if case["routine"] in ("potrs", "sytrsnp"):
        return  # inputs are factors; checked through posv/sysvnp below
    A = _dense_sym(case)
    tol = 4 * n * eps * _cond1(A)
    num = np.max(np.abs(Xg - Xo), axis=1)
    den = np.maximum(np.max(np.abs(Xo), axis=1), 1e-300)
    ratio = num / den / tol[:, None]
    assert np.all(ratio <= 1.0), f"max ratio {ratio.max()}"


@pytest.mark.parametrize("p", ["d", "s"])
@pytest.mark.parametrize("n", TEAM_N)
def test_team_factor_and_solve(p, n):
    for i, (lay, uplo) in enumerate((("C", "L"), ("C", "U"), ("R", "L"), ("R", "U"))):
        for V, nm in ((64, 130), (32, 71), (8, 13), (4, 6)):
            for rt in ("potrf", "sytrfnp", "posv", "sysvnp", "potrs", "sytrsnp"):
                for nrhs in ((0,) if rt in ("potrf", "sytrfnp") else (1, 4, 5, 24)):
                    dims = {"n": n} if nrhs == 0 else {"n": n, "nrhs": nrhs}
                    print(f"case {rt}{p} {lay}{uplo} n={n} nrhs={nrhs} V={V} nm={nm}", flush=True)
                    c = cases.make_case(rt, p, {"layout": lay, "uplo": uplo}, dims, V, nm,
                                        seed=31000 + 7 * n + i, ld_extra=i % 2)
                    so, outo = cases.run(cbkt.oracle().call, c)
                    sg, outg = cases.run(_gpu().gpu_call, c)
                    assert so == sg == 0
                    # the factor (or the untouched factor of a solve): every bit
                    assert cases.same_bits(outo[0], outg[0]), (rt, lay, uplo, V)
                    if len(outo) > 1:
                        _check_solution(c, outo[1], outg[1])


@pytest.mark.parametrize("p", ["d", "s"])
@pytest.mark.parametrize("n", TEAM_N)
def test_posv_equals_potrf_potrs_on_gpu(p, n):
    g = _gpu()
    for ldl, (sv, tf, ts) in ((False, ("posv", "potrf", "potrs")), (True, ("sysvnp", "sytrfnp", "sytrsnp"))):
        c = cases.make_case(sv, p, {"layout": "C", "uplo": "L"}, {"n": n, "nrhs": 4}, 64, 100,
                            seed=4242 + n)
        a1, b1 = c["args"][4].copy(), c["args"][6].copy()
        a2, b2 = a1.copy(), b1.copy()
        assert g.gpu_call(p, sv, "C", "L", n, 4, a1, n, b1, n, 64, 100) == 0
        assert g.gpu_call(p, tf, "C", "L", n, a2, n, 64, 100) == 0
        assert g.gpu_call(p, ts, "C", "L", n, 4, a2, n, b2, n, 64, 100) == 0
        assert cases.same_bits(a1, a2) and cases.same_bits(b1, b2)


@pytest.mark.parametrize("n", TEAM_N)
def test_team_info_codes(n):
    """Non-SPD / zero-pivot matrices: info matches the oracle (LAPACK codes),
    their B is left unmodified, and the other matrices are unaffected."""
    g = _gpu()
    for rt, f in (("posv", "potrf"), ("sysvnp", "sytrfnp")):
        c = cases.make_case(rt, "d", {"layout": "C", "uplo": "L"}, {"n": n, "nrhs": 4}, 64, 70,
                            seed=77 + n)
        ap, bp = c["args"][4], c["args"][6]
        A = cbkt.unpack(ap, 70, n, n, 64)
        bad = {3: 0, 17: n // 2, 69: n - 1}
        for mat, k in bad.items():
            if rt == "posv":
                A[mat, k, k] = -1.0
            else:
                # a zero leading minor at k: make row/col k of the LDL^T factor vanish
                A[mat, k, :k + 1] = 0.0
                A[mat, :k + 1, k] = 0.0
        c["args"][4] = cbkt.pack(A, 64)
        info_o = np.zeros(70, np.int32)
        info_g = np.zeros(70, np.int32)
        so, outo = cases.run(cbkt.oracle().call, c, info=info_o)
        sg, outg = cases.run(g.gpu_call, c, info=info_g)
        assert so == sg == 0
        assert np.array_equal(info_o, info_g), (info_o, info_g)
        # the closed form: the first failing pivot is k, info = k + 1
        assert all(info_g[m] == k + 1 for m, k in bad.items()), info_g
        assert np.count_nonzero(info_g) == len(bad)
        assert cases.same_bits_nan(outo[0], outg[0])
        Bo = cbkt.unpack(outg[1], 70, n, 4, 64)
        B0 = cbkt.unpack(bp, 70, n, 4, 64)
        for m in bad:
            assert cases.same_bits(Bo[m], B0[m])
