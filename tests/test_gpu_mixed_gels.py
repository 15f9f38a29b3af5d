"""GPU parity of dsgels_compact (mixed-precision least squares) with the
oracle's restatement (oracle_dsgels_compact, itself pinned to LAPACK dgels):
  - X: |x_gpu - x_oracle|_inf <= 20 m eps (kappa + kappa^2 ||r|| / (||A|| ||x||)) ||x||_inf;
  - iter: the same sign per matrix (refined, or sent to the FP64 path), steps
    within 2 of the oracle's (the FP32 QR sums its reductions in another order,
    which moves the step at which a correction crosses the stopping test);
  - A untouched where refinement succeeded, the FP64 QR (to rounding) where not;
  - B untouched; info through the FP64 path; iter -1 for shapes without an FP32
    kernel (V = 2, or n > 32); the FP32 kernels' buckets (m <= 16, 32, 64 rows,
    n <= 8, 16, 32 columns) at exact and padded shapes.
"""

import numpy as np
import pytest

import cbkt

pytestmark = pytest.mark.gpu


def _tall_cond(r, nm, m, n, logk):
    U = np.linalg.qr(r.standard_normal((nm, m, n)))[0]
    W = np.linalg.qr(r.standard_normal((nm, n, n)))[0]
    s = 10.0 ** (-r.uniform(0, 1, (nm, n)) * logk)
    s[:, 0], s[:, -1] = 1.0, 10.0 ** -logk
    return (U * s[:, None, :]) @ np.swapaxes(W, 1, 2)


def _run_both(layout, m, n, nrhs, V, nm, A, B):
    import gpuutil
    torch = gpuutil.torch()
    import cbk_amd
    row = layout == "R"
    ap = cbkt.pack(A, V, row=row)
    bp = cbkt.pack(B, V, row=row, pad_diag=0.0)
    ldx = nrhs if row else n
    xp = np.zeros(((nm + V - 1) // V) * n * nrhs * V)
    lda, ldb = (n if row else m), (nrhs if row else m)
    ao, bo, xo = ap.copy(), bp.copy(), xp.copy()
    it_o = np.zeros(nm, np.int32)
    inf_o = np.zeros(nm, np.int32)
    assert cbkt.oracle().call_raw("dsgels_compact", layout, "N", m, n, nrhs, ao, lda, bo, ldb, xo,
                                  ldx, it_o, inf_o, V, nm) == 0
    dA, dB, dX = (torch.from_numpy(x).cuda() for x in (ap, bp, xp))
    it_g = torch.full((nm,), -99, dtype=torch.int32, device="cuda")
    inf_g = torch.full((nm,), -99, dtype=torch.int32, device="cuda")
    st = cbk_amd.call_ext("dsgels_compact", layout, "N", m, n, nrhs, dA, lda, dB, ldb, dX, ldx,
                          it_g, inf_g, V, nm)
    assert st == 0, st
    assert cbk_amd.synchronize() == 0, cbk_amd.last_error()
    g = (dA.cpu().numpy(), dB.cpu().numpy(), dX.cpu().numpy(), it_g.cpu().numpy(), inf_g.cpu().numpy())
    o = (ao, bo, xo, it_o, inf_o)
    return ap, bp, g, o, lda, ldb, ldx


@pytest.mark.parametrize("m,n,nrhs", [(64, 32, 2), (64, 32, 1), (64, 32, 4), (32, 16, 3),
                                      (16, 8, 2), (24, 10, 2), (50, 20, 5), (64, 8, 6),
                                      (12, 5, 3), (40, 36, 2)])
@pytest.mark.parametrize("layout,V,nm", [("C", 64, 150), ("R", 8, 21), ("C", 2, 7), ("C", 32, 71)])
def test_dsgels_matches_oracle(m, n, nrhs, layout, V, nm):
    print(f"case dsgels {m}x{n} nrhs={nrhs} {layout} V={V} nm={nm}", flush=True)
    r = cbkt.rng(7000 + m + nrhs + V)
    A = r.uniform(-1, 1, (nm, m, n))
    bad = [k for k in (3, 11, 17, 140) if k < nm]
    A[bad] = _tall_cond(r, len(bad), m, n, 6)
    B = r.standard_normal((nm, m, nrhs))
    ap, bp, (ag, bg, xg, itg, infg), (ao, bo, xo, ito, info_o), lda, ldb, ldx = \
        _run_both(layout, m, n, nrhs, V, nm, A, B)
    row = layout == "R"
    assert np.array_equal(bg, bp)
    mixed = V >= 4 and n <= 32  # the FP32 buckets: m <= 64, n <= 32
    assert np.array_equal(itg < 0, ito < 0), (itg, ito)
    assert np.all(np.abs(itg - ito) <= 2), (itg, ito)
    if not mixed:
        assert np.all(itg == -1)
    else:
        assert np.all(itg[bad] == -31) and np.all(itg[[k for k in range(nm) if k not in bad]] >= 0)
    assert np.array_equal(infg, info_o) and np.all(infg == 0)
    eps = np.finfo(np.float64).eps
    Xo = cbkt.unpack(xo, nm, n, nrhs, V, ld=ldx, row=row)
    Xg = cbkt.unpack(xg, nm, n, nrhs, V, ld=ldx, row=row)
    s = np.linalg.svd(A, compute_uv=False)
    kappa = s[:, 0] / s[:, -1]
    res = B - A @ Xo
    bound = 20 * m * eps * (kappa + kappa ** 2 * np.linalg.norm(res, axis=(1, 2)) /
                            (s[:, 0] * np.linalg.norm(Xo, axis=(1, 2))))
    err = np.abs(Xg - Xo).max(axis=(1, 2)) / np.abs(Xo).max(axis=(1, 2))
    assert np.all(err <= bound), (err / bound).max()
    Fg = cbkt.unpack(ag, nm, m, n, V, ld=lda, row=row)
    Fo = cbkt.unpack(ao, nm, m, n, V, ld=lda, row=row)
    for k in range(nm):
        if itg[k] >= 0:
            assert np.array_equal(Fg[k], A[k])
        else:
            assert np.abs(Fg[k] - Fo[k]).max() <= 20 * m * eps * np.abs(A[k]).max() * kappa[k]


def test_dsgels_info_and_codes():
    import cbk_amd
    m, n, nrhs, V, nm = 32, 16, 1, 4, 8
    r = cbkt.rng(7100)
    A = r.uniform(-1, 1, (nm, m, n))
    A[5, :, 9] = 0.0
    B = r.standard_normal((nm, m, nrhs))
    _, _, (ag, bg, xg, itg, infg), (ao, bo, xo, ito, info_o), *_ = _run_both("C", m, n, nrhs, V, nm, A, B)
    assert info_o.tolist() == [0] * 5 + [10] + [0] * 2
    assert np.array_equal(infg, info_o) and itg[5] == -31
    # argument codes (nothing launched)
    z = None
    assert cbk_amd.call_ext("dsgels_compact", "C", "T", m, n, 1, z, m, z, m, z, n, z, z, V, nm) == -2
    assert cbk_amd.call_ext("dsgels_compact", "C", "N", 65, n, 1, z, 65, z, 65, z, n, z, z, V, nm) == -3
    assert cbk_amd.call_ext("dsgels_compact", "C", "N", m, m + 1, 1, z, m, z, m, z, m, z, z, V, nm) == -4
    assert cbk_amd.call_ext("dsgels_compact", "C", "N", 64, 60, 5, z, 64, z, 64, z, 60, z, z, V, nm) == -5
    assert cbk_amd.call_ext("dsgels_compact", "C", "N", m, n, 1, z, m, z, m, z, n - 1, z, z, V, nm) == -11
    assert cbk_amd.call_ext("dsgels_compact", "C", "N", m, n, 1, z, m, z, m, z, n, z, z, V, nm) == -12
    assert cbk_amd.call_ext("dsgels_compact", "C", "N", m, n, 1, z, m, z, m, z, n, 8, z, 3, nm) == -14
