"""GPU parity of the column-team QR kernel (qr_cols.hip: geqrf and the
overdetermined gels for p <= 32 rows and q + nrhs <= 16 columns, V >= 4)
against the oracle.

The kernel sums every reflector norm and dot product as an fma chain in the
reference's row order, so R, the reflectors and tau are compared bit for bit
(the zero rows that pad a matrix to 16 or 32 rows add exact zeros, which could
only turn a -0 sum into +0; random inputs give none). X comes from a
column-oriented back substitution, so it is compared within the least-squares
perturbation bound, as are the residual rows Q^T B below row q.
"""

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


# shapes the dispatcher gives this kernel (16 x 8 and 32 x 16 with nrhs <= 2 go
# to the row team kernel, test_gpu_qr_team.py)
GELS = [("d", 16, 8, 3), ("s", 16, 8, 4), ("d", 12, 5, 3), ("s", 12, 5, 3), ("d", 32, 10, 6),
        ("s", 32, 10, 6), ("d", 20, 14, 2), ("d", 7, 7, 1), ("d", 30, 1, 15)]


def _ls_check(p, Ad, B0, Bo, Bg, n, m, nm):
    eps = float(np.finfo(cbkt.DT[p]).eps)
    s = np.linalg.svd(Ad, compute_uv=False)
    kappa = s[:, 0] / s[:, -1]
    X = Bo[:, :n]
    r = B0 - Ad @ X
    xn = np.abs(X).max(axis=(1, 2))
    bound = 20 * m * eps * (kappa + kappa ** 2 * np.linalg.norm(r, axis=(1, 2)) /
                            (s[:, 0] * np.maximum(np.linalg.norm(X, axis=(1, 2)), 1e-300)))
    errx = np.abs(Bg[:, :n] - X).max(axis=(1, 2))
    assert np.all(errx <= bound * xn), (errx / (bound * xn)).max()
    bn = np.abs(B0).max(axis=(1, 2))
    errr = np.abs(Bg[:, n:] - Bo[:, n:]).max(axis=(1, 2)) if m > n else np.zeros(nm)
    assert np.all(errr <= 20 * m * eps * kappa * bn)


@pytest.mark.parametrize("p,m,n,nrhs", GELS)
@pytest.mark.parametrize("layout,trans,V", [("C", "N", 64), ("R", "N", 4), ("C", "T", 8), ("C", "N", 32)])
def test_qr_cols_gels(p, m, n, nrhs, layout, trans, V):
    nm = 3 * V + 5 if V < 64 else 140
    am, an = (m, n) if trans == "N" else (n, m)
    print(f"case gels{p} {am}x{an} nrhs={nrhs} {layout}{trans} V={V}", flush=True)
    c = cases.make_case("gels", p, {"layout": layout, "trans": trans},
                        {"m": am, "n": an, "nrhs": nrhs}, V, nm, seed=4000 + m * 7 + n + nrhs)
    so, (ao, bo, to) = cases.run(cbkt.oracle().call, c)
    sg, (ag, bg, tg) = cases.run(_gpu().gpu_call, c)
    assert so == sg == 0
    row = layout == "R"
    ua = lambda a: cbkt.unpack(a, nm, am, an, V, ld=c["args"][6], row=row)
    assert np.array_equal(_bits(ua(ag)), _bits(ua(ao))), "R and reflectors differ"
    assert np.array_equal(_bits(cbkt.unpack_vec(tg, nm, n, V)),
                          _bits(cbkt.unpack_vec(to, nm, n, V))), "tau differs"
    A = cbkt.unpack(c["args"][5], nm, am, an, V, ld=c["args"][6], row=row).astype(np.float64)
    Ad = A if trans == "N" else np.swapaxes(A, 1, 2)
    mx = max(am, an)
    unpack_b = lambda b: cbkt.unpack(b, nm, mx, nrhs, V, ld=c["args"][8], row=row).astype(np.float64)
    _ls_check(p, Ad, unpack_b(c["args"][7]), unpack_b(bo), unpack_b(bg), n, m, nm)


@pytest.mark.parametrize("p,m,n", [("d", 16, 7), ("s", 32, 15), ("d", 20, 16), ("d", 9, 4), ("s", 25, 13)])
def test_qr_cols_geqrf(p, m, n):
    V, nm = 16, 50
    print(f"case geqrf{p} {m}x{n}", flush=True)
    c = cases.make_case("geqrf", p, {"layout": "C"}, {"m": m, "n": n}, V, nm, seed=91 + m)
    so, (ao, to) = cases.run(cbkt.oracle().call, c)
    sg, (ag, tg) = cases.run(_gpu().gpu_call, c)
    assert so == sg == 0
    assert np.array_equal(_bits(cbkt.unpack(ag, nm, m, n, V)), _bits(cbkt.unpack(ao, nm, m, n, V)))
    assert np.array_equal(_bits(cbkt.unpack_vec(tg, nm, n, V)), _bits(cbkt.unpack_vec(to, nm, n, V)))


def test_qr_cols_gels_info():
    """zero columns: info = their index + 1 on the GPU and in the oracle (LAPACK
    dgels), B left as Q^T B; the other matrices solve normally"""
    V, nm, m, n, nrhs = 16, 40, 12, 6, 3
    c = cases.make_case("gels", "d", {"layout": "C", "trans": "N"}, {"m": m, "n": n, "nrhs": nrhs},
                        V, nm, seed=6)
    A = cbkt.unpack(c["args"][5], nm, m, n, V)
    A[3, :, 5] = 0.0
    A[20, :, 0] = 0.0
    c["args"][5] = cbkt.pack(A, V)
    info_o = np.zeros(nm, np.int32)
    info_g = np.zeros(nm, np.int32)
    so, (ao, bo, to) = cases.run(cbkt.oracle().call, c, info=info_o)
    sg, (ag, bg, tg) = cases.run(_gpu().gpu_call, c, info=info_g)
    assert so == sg == 0
    assert info_o[3] == 6 and info_o[20] == 1 and np.count_nonzero(info_o) == 2
    assert np.array_equal(info_o, info_g)
    Bo = cbkt.unpack(bo, nm, m, nrhs, V)
    Bg = cbkt.unpack(bg, nm, m, nrhs, V)
    for k in (3, 20):
        assert np.abs(Bo[k] - Bg[k]).max() <= 1e-12 * max(1.0, np.abs(Bo[k]).max())
