"""GPU parity of the QR team kernel (geqrf, overdetermined gels at the shapes it
serves: 16 x 8 and 32 x 16 in both precisions, 64 x 32 in FP32; selected
through the kernel-family table, since the register and lane-grid kernels come
first at some of these shapes) against the oracle. The team sums reflector norms and dot products as a butterfly, so the
comparison is to rounding:
  - R, the reflectors and tau: |gpu - oracle| <= 20 p eps ||A||_max (tau: 20 p eps);
  - X: |x_gpu - x_oracle|_inf <= 20 p eps (kappa + kappa^2 ||r|| / (||A|| ||x||)) ||x||_inf,
    the least-squares perturbation bound;
  - the residual rows of B (Q^T B beyond row q): same bound scaled by ||b||;
  - gels_info: a zero column gives info = its index + 1 and B = Q^T B, as LAPACK.
"""

import numpy as np
import pytest

import cases
import cbkt

pytestmark = pytest.mark.gpu

SHAPES = [("d", 16, 8), ("s", 16, 8), ("d", 32, 16), ("s", 32, 16), ("s", 64, 32)]


def _gpu():
    import gpuutil
    return gpuutil


def _team_run(routine, p, m, n, nrhs, V, c, **kw):
    """the case on the GPU with the team family installed for its shape"""
    import cbk_amd
    F = cbk_amd.FAMILY
    assert cbk_amd.set_kernel_family(p, routine, m, n, nrhs, V, F["team"]) == 0
    try:
        return cases.run(_gpu().gpu_call, c, **kw)
    finally:
        cbk_amd.set_kernel_family(p, routine, m, n, nrhs, V, F["default"])


@pytest.mark.parametrize("p,m,n", SHAPES)
@pytest.mark.parametrize("layout,trans,V", [("C", "N", 64), ("R", "N", 64), ("C", "T", 64),
                                           ("C", "N", 32)])
def test_qr_team_gels(p, m, n, layout, trans, V):
    nm = 140 if V == 64 else 3 * V + 7
    # trans: op(A) = A^T with A n x m (m < n as stored) is the same overdetermined problem
    am, an = (m, n) if trans == "N" else (n, m)
    for nrhs in (1, 2):
        print(f"case gels{p} {am}x{an} {layout}{trans} nrhs={nrhs}", flush=True)
        c = cases.make_case("gels", p, {"layout": layout, "trans": trans},
                            {"m": am, "n": an, "nrhs": nrhs}, V, nm, seed=3000 + m + nrhs)
        so, (ao, bo, to) = cases.run(cbkt.oracle().call, c)
        sg, (ag, bg, tg) = _team_run("gels", p, m, n, nrhs, V, c)
        assert so == sg == 0
        eps = float(np.finfo(cbkt.DT[p]).eps)
        row = layout == "R"
        A = cbkt.unpack(c["args"][5], nm, am, an, V, ld=c["args"][6], row=row).astype(np.float64)
        Ad = A if trans == "N" else np.swapaxes(A, 1, 2)   # the m x n overdetermined operator
        amax = np.abs(A).max(axis=(1, 2))
        Ao = cbkt.unpack(ao, nm, am, an, V, ld=c["args"][6], row=row).astype(np.float64)
        Ag = cbkt.unpack(ag, nm, am, an, V, ld=c["args"][6], row=row).astype(np.float64)
        assert np.all(np.abs(Ag - Ao).max(axis=(1, 2)) <= 20 * m * eps * amax)
        To = cbkt.unpack_vec(to, nm, n, V).astype(np.float64)
        Tg = cbkt.unpack_vec(tg, nm, n, V).astype(np.float64)
        assert np.all(np.abs(Tg - To) <= 20 * m * eps)
        mx = max(am, an)
        Bo = cbkt.unpack(bo, nm, mx, nrhs, V, ld=c["args"][8], row=row).astype(np.float64)
        Bg = cbkt.unpack(bg, nm, mx, nrhs, V, ld=c["args"][8], row=row).astype(np.float64)
        B0 = cbkt.unpack(c["args"][7], nm, mx, nrhs, V, ld=c["args"][8], row=row).astype(np.float64)
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


@pytest.mark.parametrize("p,m,n", [("d", 16, 8), ("s", 32, 16)])
def test_qr_team_geqrf(p, m, n):
    V, nm = 16, 50
    c = cases.make_case("geqrf", p, {"layout": "C"}, {"m": m, "n": n}, V, nm, seed=77 + m)
    so, (ao, to) = cases.run(cbkt.oracle().call, c)
    sg, (ag, tg) = _team_run("geqrf", p, m, n, 0, V, c)
    assert so == sg == 0
    eps = float(np.finfo(cbkt.DT[p]).eps)
    A = cbkt.unpack(c["args"][3], nm, m, n, V).astype(np.float64)
    amax = np.abs(A).max(axis=(1, 2))
    d = np.abs(cbkt.unpack(ag, nm, m, n, V).astype(np.float64) -
               cbkt.unpack(ao, nm, m, n, V).astype(np.float64)).max(axis=(1, 2))
    assert np.all(d <= 20 * m * eps * amax)
    assert np.all(np.abs(cbkt.unpack_vec(tg, nm, n, V) - cbkt.unpack_vec(to, nm, n, V)) <= 20 * m * eps)


def test_qr_team_gels_info():
    """a zero column: info = its index + 1 on the GPU and in the oracle (LAPACK
    dgels), B left as Q^T B; the other matrices solve normally"""
    V, nm, m, n = 64, 70, 16, 8
    c = cases.make_case("gels", "d", {"layout": "C", "trans": "N"}, {"m": m, "n": n, "nrhs": 2},
                        V, nm, seed=5)
    A = cbkt.unpack(c["args"][5], nm, m, n, V)
    A[10, :, 3] = 0.0
    A[40, :, 0] = 0.0
    c["args"][5] = cbkt.pack(A, V)
    info_o = np.zeros(nm, np.int32)
    info_g = np.zeros(nm, np.int32)
    so, (ao, bo, to) = cases.run(cbkt.oracle().call, c, info=info_o)
    sg, (ag, bg, tg) = _team_run("gels", "d", m, n, 2, V, c, info=info_g)
    assert so == sg == 0
    assert info_o[10] == 4 and info_o[40] == 1 and np.count_nonzero(info_o) == 2
    assert np.array_equal(info_o, info_g)
    Bo = cbkt.unpack(bo, nm, m, 2, V)
    Bg = cbkt.unpack(bg, nm, m, 2, V)
    for k in (10, 40):
        assert np.abs(Bo[k] - Bg[k]).max() <= 1e-12 * max(1.0, np.abs(Bo[k]).max())
