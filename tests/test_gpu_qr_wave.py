"""GPU parity of gels_compact on the shapes where neither QR team kernel
applies (the tall view p x q with p <= 64 and q + nrhs <= 64: more than 32
rows, more than 16 columns of [A | B], or V < 4), through the default dispatch
(the lane-grid kernels: two problems per wavefront on 4 x 8 lanes each for
q <= 32, one on 8 x 8 lanes above, nrhs <= 8) and through the wavefront kernel
(one matrix per wavefront, lane c owning column c of [A | B]) forced by the
kernel-family table, against the oracle. Sums are in another order than the
reference's (per-lane and cross-lane partial sums), so the comparison is to
the least-squares perturbation bound, as for the team kernel:
  - R, the reflectors and tau: |gpu - oracle| <= 20 p eps ||A||_max (tau: 20 p eps);
  - X: |x_gpu - x_oracle|_inf <= 20 p eps (kappa + kappa^2 ||r|| / (||A|| ||x||)) ||x||_inf;
  - the rows of Q^T B below q: 20 p eps kappa ||b||;
  - gels_info: a zero column gives info = its index + 1 and B = Q^T B.
"""

import numpy as np
import pytest

import cases
import cbkt

pytestmark = pytest.mark.gpu

SHAPES = [("d", 64, 32, 2), ("d", 64, 32, 1), ("d", 24, 14, 3), ("s", 40, 20, 4),
          ("d", 64, 60, 4), ("s", 33, 1, 1), ("d", 48, 30, 8)]


def _gpu():
    import gpuutil
    return gpuutil


@pytest.mark.parametrize("family", ["default", "wave"])
@pytest.mark.parametrize("p,m,n,nrhs", SHAPES)
@pytest.mark.parametrize("layout,trans,V,nm", [("C", "N", 64, 130), ("R", "N", 16, 37),
                                               ("C", "T", 2, 9), ("C", "N", 32, 71)])
def test_qr_wave_gels(p, m, n, nrhs, layout, trans, V, nm, family):
    import cbk_amd
    am, an = (m, n) if trans == "N" else (n, m)
    print(f"case gels-{family} {p} {am}x{an} nrhs={nrhs} {layout}{trans} V={V} nm={nm}", flush=True)
    c = cases.make_case("gels", p, {"layout": layout, "trans": trans},
                        {"m": am, "n": an, "nrhs": nrhs}, V, nm, seed=5000 + m + n + nrhs)
    so, (ao, bo, to) = cases.run(cbkt.oracle().call, c)
    F = cbk_amd.FAMILY
    assert cbk_amd.set_kernel_family(p, "gels", m, n, nrhs, V, F[family]) == 0
    try:
        sg, (ag, bg, tg) = cases.run(_gpu().gpu_call, c)
    finally:
        cbk_amd.set_kernel_family(p, "gels", m, n, nrhs, V, F["default"])
    assert so == sg == 0
    eps = float(np.finfo(cbkt.DT[p]).eps)
    row = layout == "R"
    A = cbkt.unpack(c["args"][5], nm, am, an, V, ld=c["args"][6], row=row).astype(np.float64)
    Ad = A if trans == "N" else np.swapaxes(A, 1, 2)
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
    if m > n:
        errr = np.abs(Bg[:, n:] - Bo[:, n:]).max(axis=(1, 2))
        assert np.all(errr <= 20 * m * eps * kappa * np.abs(B0).max(axis=(1, 2)))
    # the padded slots of the last group keep their identity / zero content
    if nm % V:
        assert cases.same_bits(cbkt.unpack(ag, nm, am, an, V, ld=c["args"][6], row=row, all_slots=True)[nm:],
                               cbkt.unpack(ao, nm, am, an, V, ld=c["args"][6], row=row, all_slots=True)[nm:])


@pytest.mark.parametrize("n,zcol", [(32, 17), (48, 40)])
def test_qr_wave_gels_info(n, zcol):
    V, nm, m = 64, 70, 64
    c = cases.make_case("gels", "d", {"layout": "C", "trans": "N"}, {"m": m, "n": n, "nrhs": 2},
                        V, nm, seed=77)
    A = cbkt.unpack(c["args"][5], nm, m, n, V)
    A[10, :, zcol] = 0.0
    A[40, :, 0] = 0.0
    c["args"][5] = cbkt.pack(A, V)
    info_o = np.zeros(nm, np.int32)
    info_g = np.zeros(nm, np.int32)
    so, (ao, bo, to) = cases.run(cbkt.oracle().call, c, info=info_o)
    sg, (ag, bg, tg) = cases.run(_gpu().gpu_call, c, info=info_g)
    assert so == sg == 0
    assert info_o[10] == zcol + 1 and info_o[40] == 1 and np.count_nonzero(info_o) == 2
    assert np.array_equal(info_o, info_g)
    Bo = cbkt.unpack(bo, nm, m, 2, V)
    Bg = cbkt.unpack(bg, nm, m, 2, V)
    for k in (10, 40):
        assert np.abs(Bo[k] - Bg[k]).max() <= 1e-12 * max(1.0, np.abs(Bo[k]).max())
