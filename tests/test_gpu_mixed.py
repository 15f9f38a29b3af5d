"""Mixed-precision dposv (dsposv_compact) and sysvnp (dssysvnp_compact) on the
GPU against the FP64 oracle.

Checks, per matrix:
  - backward error max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)
    <= 1e-14, computed in FP64 on the host from the dense input;
  - X agrees with the oracle's dsposv (LAPACK dsposv restated) within
    4 n eps cond_1(A) per column;
  - refinement succeeded (iter >= 0) and A is untouched, bit for bit;
  - forced fallback (cond ~ 1e12 > 1/eps_32): iter < 0, A holds the FP64
    Cholesky factor bit-identical to the oracle's potrf, X matches dposv;
  - a non-SPD matrix: info matches the oracle's FP64 dpotrf code;
  - the same for dssysvnp_compact on symmetric indefinite matrices (the FP64
    LDL^T factor bit-identical to the oracle's sytrfnp, info its first zero
    pivot);
  - every ITER code from a construction that fixes it (-2: B, A entry beyond
    FLT_MAX; -3: a diagonal that is 0 in FP32, for sysvnp also a D whose FP32
    reciprocal overflows), equal to the oracle's, X then bit-identical to the
    oracle's FP64 solve.
Both the fused kernels (n in 1..64, any nrhs in blocks; the quad kernel at
17..24, the team kernel at the other orders up to 32, an order padded with the
identity up to its kernel's: 24, or 16 / 24 / 32; the row-per-lane kernel at
33..64, its failures on the FP64 grid kernel) and the
workspace lane kernel (other shapes, or the lane policy) are exercised.
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


def run_dsposv(p_call, A, B, V, uplo="L", layout="C", name="dsposv_compact"):
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
        st = cbkt.oracle().call_raw(name, layout, uplo, n, nrhs, ap, n, bp, ldb, xp,
                                    ldb, None, 0, iters, info, V, nm)
        return st, ap, xp, iters, info
    import cbk_amd
    import gpuutil
    torch = gpuutil.torch()
    ws = cbk_amd.dsposv_worksize(layout, uplo, n, nrhs, n, ldb, ldb, V, nm)
    dA, dB, dX = (gpuutil.to_dev(x) for x in (ap, bp, xp))
    dI, dN = gpuutil.to_dev(iters), gpuutil.to_dev(info)
    dW = torch.empty(max(ws, 8), dtype=torch.uint8, device="cuda") if ws else None
    st = cbk_amd.call_ext(name, layout, uplo, n, nrhs, dA, n, dB, ldb, dX, ldb, dW, ws,
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
    (20, 5, 8, "L", "C", False),    # orders between 16, 24 and 32: padded with the identity
    (17, 3, 64, "U", "R", False),   # (17..23 on the quad kernel, the others on the team kernel)
    (23, 24, 16, "L", "C", False),
    (28, 4, 64, "L", "C", False),
    (12, 3, 8, "U", "R", False),
    (20, 5, 8, "L", "C", True),     # the same order through the lane kernel
    (40, 4, 64, "L", "C", False),   # 33..64: one matrix per wavefront, a row per lane
    (64, 3, 8, "U", "R", False),
    (33, 9, 16, "L", "C", False),
    (48, 1, 4, "U", "C", False),
    (7, 2, 2, "U", "R", False),     # padded up to 16
    (7, 2, 2, "U", "R", True),
    (1, 3, 4, "L", "C", False),
    (32, 4, 32, "L", "C", False),   # V = 32: the fused kernel's matrices span two per group line
    (24, 2, 32, "U", "R", False),
    (32, 9, 64, "L", "C", False),   # nrhs > 4: the fused kernel in blocks of 4 columns
    (24, 24, 16, "U", "C", False),
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


@pytest.mark.parametrize("n,V,lane", [(32, 64, False), (32, 64, True), (24, 8, False),
                                      (28, 8, False), (14, 8, False), (20, 8, False),
                                      (40, 8, False), (64, 64, False)])
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


def sym_indefinite_logspace(r, nm, n, logk):
    """Q diag(+/-lambda) Q^T, log10 lambda uniform on [0, logk]: indefinite,
    with (generically) an unpivoted LDL^T"""
    Q = np.linalg.qr(r.standard_normal((nm, n, n)))[0]
    lam = 10.0 ** (r.uniform(0, 1, (nm, n)) * logk)
    lam = lam * np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    A = (Q * lam[:, None, :]) @ np.swapaxes(Q, 1, 2)
    return 0.5 * (A + np.swapaxes(A, 1, 2))


@pytest.mark.parametrize("n,nrhs,V,uplo,layout,lane", [
    (24, 24, 64, "L", "C", False),  # the sysvnp benchmark shape, fused kernel
    (32, 4, 64, "L", "C", False),
    (16, 1, 4, "U", "R", False),
    (24, 24, 64, "L", "C", True),   # the same through the workspace lane kernel
    (20, 5, 8, "L", "C", False),    # padded with the identity up to 24
    (30, 4, 16, "U", "R", False),   # up to 32
    (40, 4, 64, "L", "C", False),   # 33..64: a row per lane
    (64, 5, 8, "U", "R", False),
    (20, 5, 8, "L", "C", True),     # the lane kernel
    (24, 6, 32, "U", "C", False),
])
def test_dssysvnp_refines_to_fp64_accuracy(n, nrhs, V, uplo, layout, lane):
    import cbk_amd
    r = cbkt.rng(300 + n + nrhs)
    nm = 2 * V + 3
    A = cbkt.gen_sym_dominant(r, nm, n)
    B = r.standard_normal((nm, n, nrhs))
    print(f"case dssysvnp n={n} nrhs={nrhs} V={V} {layout}{uplo} lane={lane}", flush=True)
    prev = cbk_amd.set_policy(cbk_amd.POLICY_LANE if lane else cbk_amd.POLICY_AUTO)
    try:
        st, ap, xp, iters, info = run_dsposv("gpu", A, B, V, uplo, layout, "dssysvnp_compact")
    finally:
        cbk_amd.set_policy(prev)
    assert st == 0
    assert np.all(info == 0)
    assert np.all(iters >= 0), iters
    row = layout == "R"
    X = cbkt.unpack(xp, nm, n, nrhs, V, row=row)
    eta = backward_error(A, X, B)
    assert np.all(eta <= 1e-14), eta.max()
    Aref = np.where(np.eye(n, dtype=bool), A, np.tril(A) if uplo == "L" else np.triu(A))
    assert np.array_equal(ap, cbkt.pack(Aref, V, row=row))
    so, _, xo, io, no = run_dsposv("oracle", A, B, V, uplo, layout, "dssysvnp_compact")
    assert so == 0 and np.all(io >= 0)
    Xo = cbkt.unpack(xo, nm, n, nrhs, V, row=row)
    tol = 4 * n * 2.0 ** -53 * cond1(A)
    ratio = np.abs(X - Xo).max(axis=1) / np.abs(Xo).max(axis=1) / tol[:, None]
    assert np.all(ratio <= 1.0), ratio.max()


@pytest.mark.parametrize("n,V,lane", [(24, 64, False), (24, 64, True), (32, 8, False),
                                      (21, 8, False), (27, 8, False), (48, 8, False)])
def test_dssysvnp_fallback_and_info(n, V, lane):
    import cbk_amd
    r = cbkt.rng(17 + n)
    nm = V + 3
    nrhs = 3
    A = cbkt.gen_sym_dominant(r, nm, n)
    A[[2, 6, nm - 1]] = sym_indefinite_logspace(r, 3, n, 11)  # FP32 cannot refine these
    A[8, :, 0] = A[8, 0, :] = 0.0                             # D(1) = 0 exactly: FP64 fails too
    A[8, 0, 0] = 0.0
    B = r.standard_normal((nm, n, nrhs))
    print(f"case dssysvnp fallback n={n} V={V} lane={lane}", flush=True)
    prev = cbk_amd.set_policy(cbk_amd.POLICY_LANE if lane else cbk_amd.POLICY_AUTO)
    try:
        st, ap, xp, iters, info = run_dsposv("gpu", A, B, V, name="dssysvnp_compact")
    finally:
        cbk_amd.set_policy(prev)
    so, apo, xo, io, no = run_dsposv("oracle", A, B, V, name="dssysvnp_compact")
    assert st == so == 0
    assert np.array_equal(info, no), (info, no)
    assert info[8] == 1 and np.count_nonzero(info) == 1
    for m in (2, 6, nm - 1, 8):
        assert iters[m] < 0 and io[m] < 0, (m, iters[m], io[m])
    ok = [m for m in range(nm) if m not in (2, 6, nm - 1, 8)]
    assert np.all(iters[ok] >= 0)
    Ag = cbkt.unpack(ap, nm, n, n, V)
    Ao = cbkt.unpack(apo, nm, n, n, V)
    X = cbkt.unpack(xp, nm, n, nrhs, V)
    Xo = cbkt.unpack(xo, nm, n, nrhs, V)
    for m in (2, 6, nm - 1):
        # the FP64 LDL^T factor bit for bit, and the FP64 solve
        assert np.array_equal(np.tril(Ag[m]), np.tril(Ao[m]))
        tol = 4 * n * 2.0 ** -53 * cond1(A[m:m + 1])[0]
        assert np.abs(X[m] - Xo[m]).max() / np.abs(Xo[m]).max() <= tol


def iter_code_cases(n, ldl):
    """matrices whose ITER code the construction fixes (as the oracle's own
    test of these codes): 1, 3: -2 from B; 2: -2 from A; 4: -3 (a diagonal
    1e-50, zero in FP32); 5: -3 with FP64 info 3; 6 (sysvnp): -3 from a D
    whose FP32 reciprocal overflows; 0, 7: refined"""
    r = cbkt.rng(33 + n)
    nm = 8
    A = cbkt.gen_spd(r, nm, n)
    B = r.standard_normal((nm, n, 2))
    B[1, 2, 1] = 1e39
    A[2, 4, 1] = A[2, 1, 4] = 1e39
    A[2, 4, 4] = 1e80
    A[3, 1, 1] = 1e39
    B[3, 0, 0] = -1e39
    A[4] = np.diag(np.full(n, 2.0))
    A[4, 0, 0] = 1e-50
    A[5] = np.diag(np.arange(1.0, n + 1.0))
    A[5, 2, 2] = 0.0
    if ldl:
        A[6] = np.eye(n) + 0.5 * (np.eye(n, k=1) + np.eye(n, k=-1))
        A[6, 0, 0] = 1e-40
    return A, B


@pytest.mark.parametrize("name", ["dsposv_compact", "dssysvnp_compact"])
@pytest.mark.parametrize("n,V,lane", [(16, 8, False), (32, 4, False), (6, 8, False),
                                      (6, 8, True), (16, 8, True), (11, 8, False),
                                      (19, 8, False), (29, 4, False), (40, 8, False),
                                      (64, 4, False)])
def test_mixed_iter_codes_match_oracle(name, n, V, lane):
    import cbk_amd
    ldl = name == "dssysvnp_compact"
    A, B = iter_code_cases(n, ldl)
    nm = A.shape[0]
    print(f"case {name} iter codes n={n} V={V} lane={lane}", flush=True)
    prev = cbk_amd.set_policy(cbk_amd.POLICY_LANE if lane else cbk_amd.POLICY_AUTO)
    try:
        st, ap, xp, iters, info = run_dsposv("gpu", A, B, V, name=name)
    finally:
        cbk_amd.set_policy(prev)
    so, apo, xo, io, no = run_dsposv("oracle", A, B, V, name=name)
    assert st == so == 0
    assert np.array_equal(info, no), (info, no)
    assert info[5] == 3
    fixed = [1, 2, 3, 4, 5] + ([6] if ldl else [])
    assert np.array_equal(iters[fixed], io[fixed]), (iters, io)
    assert iters[1] == iters[2] == iters[3] == -2 and iters[4] == -3 and iters[5] == -3
    assert np.all(iters[[0, 7]] >= 0)
    X = cbkt.unpack(xp, nm, n, 2, V)
    Xo = cbkt.unpack(xo, nm, n, 2, V)
    Ag = cbkt.unpack(ap, nm, n, n, V)
    Ao = cbkt.unpack(apo, nm, n, n, V)
    for m in [1, 2, 3, 4] + ([6] if ldl else []):
        # the FP64 path ran: factor bit-identical, X within the solve's rounding
        assert np.array_equal(np.tril(Ag[m]), np.tril(Ao[m])), m
        assert np.allclose(X[m], Xo[m], rtol=1e-12, atol=1e-12 * np.abs(Xo[m]).max()), m


@pytest.mark.parametrize("name", ["dsposv_compact", "dssysvnp_compact"])
@pytest.mark.parametrize("n,nrhs,V,uplo,layout", [(20, 3, 8, "L", "C"), (30, 5, 16, "U", "C"),
                                                  (40, 4, 8, "L", "C"), (64, 2, 4, "U", "R")])
def test_mixed_padded_leading_dimensions(name, n, nrhs, V, uplo, layout):
    """ldap, ldbp, ldxp beyond the order: the same X as the oracle's to its
    tolerance, A untouched on success, and the padding of A and X untouched
    (the fused kernels address rows and columns through the strides)"""
    import cbk_amd
    import gpuutil
    torch = gpuutil.torch()
    ldl = name == "dssysvnp_compact"
    r = cbkt.rng(700 + n + nrhs)
    nm = 2 * V + 3
    A = sym_indefinite_logspace(r, nm, n, 3) if ldl else spd_logspace(r, nm, n)
    B = r.standard_normal((nm, n, nrhs))
    row = layout == "R"
    lda = n + 3
    ldb = (nrhs if row else n) + 2
    Aref = np.where(np.eye(n, dtype=bool), A, np.tril(A) if uplo == "L" else np.triu(A))
    ap = cbkt.pack(Aref, V, ld=lda, row=row)
    bp = cbkt.pack(B, V, ld=ldb, row=row, pad_diag=0.0)
    xp = np.full_like(bp, cbkt.SENTINEL)
    print(f"case {name} padded ld n={n} nrhs={nrhs} V={V} {layout}{uplo}", flush=True)
    dA, dB, dX = (gpuutil.to_dev(x) for x in (ap, bp, xp))
    dI = gpuutil.to_dev(np.zeros(nm, np.int32))
    dN = gpuutil.to_dev(np.zeros(nm, np.int32))
    ws = cbk_amd.dsposv_worksize(layout, uplo, n, nrhs, lda, ldb, ldb, V, nm)
    dW = torch.empty(max(ws, 8), dtype=torch.uint8, device="cuda") if ws else None
    st = cbk_amd.call_ext(name, layout, uplo, n, nrhs, dA, lda, dB, ldb, dX, ldb, dW, ws, dI, dN, V, nm)
    assert cbk_amd.synchronize() == 0, cbk_amd.last_error()
    assert st == 0
    ag, xg, it, info = dA.cpu().numpy(), dX.cpu().numpy(), dI.cpu().numpy(), dN.cpu().numpy()
    assert np.all(info == 0) and np.all(it >= 0), (it, info)
    assert np.array_equal(ag, ap)                       # A (and its padding) untouched
    X = cbkt.unpack(xg, nm, n, nrhs, V, ld=ldb, row=row)
    # X's padding rows / columns keep their sentinel
    Xall = cbkt.unpack(xg, nm, ldb if not row else n, nrhs if not row else ldb, V, ld=ldb, row=row)
    if row:
        assert np.all(Xall[:, :, nrhs:] == cbkt.SENTINEL)
    else:
        assert np.all(Xall[:, n:, :] == cbkt.SENTINEL)
    so, _, xo, io, no = run_dsposv("oracle", A, B, V, uplo, layout, name=name)
    assert so == 0 and np.all(io >= 0)
    Xo = cbkt.unpack(xo, nm, n, nrhs, V, row=row)
    tol = 4 * n * 2.0 ** -53 * cond1(A)
    ratio = np.abs(X - Xo).max(axis=1) / np.abs(Xo).max(axis=1) / tol[:, None]
    assert np.all(ratio <= 1.0), ratio.max()
