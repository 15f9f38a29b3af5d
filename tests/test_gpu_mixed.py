"""Mixed-precision dposv (dsposv_compact) on the GPU against the FP64 oracle.

Checks, per matrix:
  - backward error max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)
    <= 1e-14, computed in FP64 on the host from the dense input;
  - X agrees with the oracle's dsposv (LAPACK dsposv restated) within
    4 n eps cond_1(A) per column;
  - refinement succeeded (iter >= 0) and A is untouched, bit for bit;
  - forced fallback (cond ~ 1e12 > 1/eps_32): iter < 0, A holds the FP64
    Cholesky factor bit-identical to the oracle's potrf, X matches dposv;
  - a non-SPD matrix: info matches the oracle's FP64 dpotrf code.
Both the fused team kernel (n in {16,24,32}, nrhs <= 4) and the workspace lane
kernel (other shapes, or the lane policy) are exercised.
"""

import numpy as np
import pytest

import cbkt

pytestmark = pytest.mark.gpu


def spd_logspace(r, nm, n, kmax_choices=(2, 4, 6)):
    G = r.standard_normal((nm, n, n))
    Q, Rr = np.linalg.qr(G)
    Q = Q * np.sign(np.diagonal(Rr, axis1=1, axis2=2))[:, None, :]
    k = np.array(kmax_choices, dtype=float)[r.integers(0, len(kmax_choices), nm)]
    lam = 10.0 ** (r.uniform(0, 1, (nm, n)) * k[:, None])
    A = (Q * lam[:, None, :]) @ np.swapaxes(Q, 1, 2)
    return 0.5 * (A + np.swapaxes(A, 1, 2))


def run_dsposv(p_call, A, B, V, uplo="L", layout="C"):
    nm, n, _ = A.shape
    nrhs = B.shape[2]
    row = layout == "R"
    Aref = np.tril(A) if uplo == "L" else np.triu(A)  # only the named triangle is referenced
    ap = cbkt.pack(np.where(np.eye(n, dtype=bool), A, Aref), V, row=row)
    ldb = nrhs if row else n
    bp = cbkt.pack(B, V, row=row, pad_diag=0.0)
    xp = np.zeros_like(bp)
    iters = np.zeros(nm, np.int32)
    info = np.zeros(nm, np.int32)
    if p_call == "oracle":
        st = cbkt.oracle().call_raw("dsposv_compact", layout, uplo, n, nrhs, ap, n, bp, ldb, xp,
                                    ldb, None, 0, iters, info, V, nm)
        return st, ap, xp, iters, info
    import cbk_amd
    import gpuutil
    torch = gpuutil.torch()
    ws = cbk_amd.dsposv_worksize(layout, uplo, n, nrhs, n, ldb, ldb, V, nm)
    dA, dB, dX = (gpuutil.to_dev(x) for x in (ap, bp, xp))
    dI, dN = gpuutil.to_dev(iters), gpuutil.to_dev(info)
    dW = torch.empty(max(ws, 8), dtype=torch.uint8, device="cuda") if ws else None
    st = cbk_amd.call_ext("dsposv_compact", layout, uplo, n, nrhs, dA, n, dB, ldb, dX, ldb, dW, ws,
                          dI, dN, V, nm)
    assert cbk_amd.synchronize() == 0, cbk_amd.last_error()
    return st, dA.cpu().numpy(), dX.cpu().numpy(), dI.cpu().numpy(), dN.cpu().numpy()


def backward_error(A, X, B):
    R = B - A @ X
    an = np.abs(A).sum(axis=2).max(axis=1)
    return (np.abs(R).max(axis=1) / (an[:, None] * np.abs(X).max(axis=1) + np.abs(B).max(axis=1))).max(axis=1)


def cond1(A):
    return np.linalg.norm(A, 1, axis=(1, 2)) * np.linalg.norm(np.linalg.inv(A), 1, axis=(1, 2))


@pytest.mark.parametrize("n,nrhs,V,uplo,layout,lane", [
    (32, 4, 64, "L", "C", False),   # the fused kernel at the benchmark shape
    (24, 3, 8, "U", "C", False),
    (16, 1, 4, "L", "R", False),
    (32, 4, 64, "L", "C", True),    # the same shape through the workspace lane kernel
    (20, 5, 8, "L", "C", False),    # an order only the lane kernel serves
    (7, 2, 2, "U", "R", False),
    (32, 4, 32, "L", "C", False),   # V = 32: the fused kernel's matrices span two per group line
    (24, 2, 32, "U", "R", False),
])
def test_dsposv_refines_to_fp64_accuracy(n, nrhs, V, uplo, layout, lane):
    import cbk_amd
    r = cbkt.rng(100 + n + nrhs)
    nm = 3 * V + 5
    A = spd_logspace(r, nm, n)
    B = r.standard_normal((nm, n, nrhs))
    print(f"case dsposv n={n} nrhs={nrhs} V={V} {layout}{uplo} lane={lane}", flush=True)
    prev = cbk_amd.set_policy(cbk_amd.POLICY_LANE if lane else cbk_amd.POLICY_AUTO)
    try:
        st, ap, xp, iters, info = run_dsposv("gpu", A, B, V, uplo, layout)
    finally:
        cbk_amd.set_policy(prev)
    assert st == 0
    assert np.all(info == 0)
    assert np.all(iters >= 0), iters
    row = layout == "R"
    X = cbkt.unpack(xp, nm, n, nrhs, V, row=row)
    eta = backward_error(A, X, B)
    assert np.all(eta <= 1e-14), eta.max()
    # A untouched on success
    Aref = np.where(np.eye(n, dtype=bool), A, np.tril(A) if uplo == "L" else np.triu(A))
    assert np.array_equal(ap, cbkt.pack(Aref, V, row=row))
    # against the oracle's dsposv
    so, _, xo, io, no = run_dsposv("oracle", A, B, V, uplo, layout)
    assert so == 0 and np.all(io >= 0)
    Xo = cbkt.unpack(xo, nm, n, nrhs, V, row=row)
    tol = 4 * n * 2.0 ** -53 * cond1(A)
    ratio = np.abs(X - Xo).max(axis=1) / np.abs(Xo).max(axis=1) / tol[:, None]
    assert np.all(ratio <= 1.0), ratio.max()


@pytest.mark.parametrize("n,V,lane", [(32, 64, False), (32, 64, True), (24, 8, False)])
def test_dsposv_fallback_and_info(n, V, lane):
    import cbk_amd
    r = cbkt.rng(7 + n)
    nm = V + 3
    nrhs = 2
    A = spd_logspace(r, nm, n, kmax_choices=(2,))
    bad = spd_logspace(r, 3, n, kmax_choices=(12,))      # cond ~ 1e12: FP32 cannot refine it
    A[[1, 5, nm - 1]] = bad
    A[9] = A[9] - 2.0 * np.eye(n) * np.abs(A[9]).max()    # not SPD: FP64 dpotrf fails too
    B = r.standard_normal((nm, n, nrhs))
    print(f"case dsposv fallback n={n} V={V} lane={lane}", flush=True)
    prev = cbk_amd.set_policy(cbk_amd.POLICY_LANE if lane else cbk_amd.POLICY_AUTO)
    try:
        st, ap, xp, iters, info = run_dsposv("gpu", A, B, V)
    finally:
        cbk_amd.set_policy(prev)
    so, apo, xo, io, no = run_dsposv("oracle", A, B, V)
    assert st == so == 0
    assert np.array_equal(info, no), (info, no)
    assert info[9] > 0 and np.count_nonzero(info) == 1
    for m in (1, 5, nm - 1, 9):
        assert iters[m] < 0, (m, iters[m])
    ok = [m for m in range(nm) if m not in (1, 5, nm - 1, 9)]
    assert np.all(iters[ok] >= 0)
    # fallback matrices: A holds the FP64 factor, bit-identical to the oracle's
    Ag = cbkt.unpack(ap, nm, n, n, V)
    Ao = cbkt.unpack(apo, nm, n, n, V)
    for m in (1, 5, nm - 1):
        assert np.array_equal(np.tril(Ag[m]), np.tril(Ao[m]))
    X = cbkt.unpack(xp, nm, n, nrhs, V)
    Xo = cbkt.unpack(xo, nm, n, nrhs, V)
    for m in (1, 5, nm - 1):
        tol = 4 * n * 2.0 ** -53 * cond1(A[m:m + 1])[0]
        assert np.abs(X[m] - Xo[m]).max() / np.abs(Xo[m]).max() <= tol
