"""GPU parity of the mixed-precision uneven batch (dsgels_batch) against the
oracle's restatement (oracle_dsgels_batch) and against the library's own
FP64 batch (dgels_batch):

  - refined problems (iter >= 0 on both): X within the least-squares
    perturbation bound of the oracle's X, A untouched bit for bit;
  - cond 1e12 problems (-31) and shapes outside the FP32 path (-1): the same
    code as the oracle, and A and X bit-identical to dgels_batch on the same
    problem (the fallback runs those kernels, X := B first); info as theirs;
  - a K5-distributed batch (m ~ U[4,64], n ~ U[2,m], nrhs ~ U[1,8]) converges
    to FP64 backward error per column, the least-squares ratio of LAPACK's
    own tests (dqrt17): ||A^T r||_2 / (||A||_F (||A||_F ||x||_2 + ||b||_2))
    <= 1e-15, which an FP64 Householder solve meets with eps-level values
    (and which, unlike ||A^T r|| / (||A|| ||r||), stays meaningful when the
    residual is small);
  - replays of the recorded graph give the same bits; new buffers re-record.
"""

import numpy as np
import pytest

import batchcase
import cbk_amd
import cbkt

pytestmark = pytest.mark.gpu


def _t():
    import gpuutil
    return gpuutil.torch()


def _mixed(desc, ap, bp, trans="N", reps=1):
    torch = _t()
    plan = cbk_amd.GelsBatchPlan("M", trans, desc)
    dA0 = torch.from_numpy(ap).cuda()
    dB = torch.from_numpy(bp).cuda()
    work = torch.empty(max(plan.worksize, 1), dtype=torch.uint8, device="cuda")
    outs = []
    for _ in range(reps):
        dA = dA0.clone()
        dX = torch.full_like(dB, 7.0)
        it = torch.full((len(desc),), -99, dtype=torch.int32, device="cuda")
        info = torch.full((len(desc),), -99, dtype=torch.int32, device="cuda")
        st = plan.run_mixed(dA, dB, dX, it, info, work)
        assert st == 0, (st, cbk_amd.last_error())
        assert cbk_amd.synchronize() == 0, cbk_amd.last_error()
        outs.append((dA.cpu().numpy(), dX.cpu().numpy(), it.cpu().numpy(), info.cpu().numpy()))
    return plan, outs


def _plain(desc, ap, bp, trans="N"):
    torch = _t()
    plan = cbk_amd.GelsBatchPlan("D", trans, desc)
    dA = torch.from_numpy(ap).cuda()
    dB = torch.from_numpy(bp).cuda()
    info = torch.full((len(desc),), -9, dtype=torch.int32, device="cuda")
    work = torch.empty(max(plan.worksize, 1), dtype=torch.uint8, device="cuda")
    assert plan.run(dA, dB, info, work) == 0
    assert cbk_amd.synchronize() == 0
    return dA.cpu().numpy(), dB.cpu().numpy(), info.cpu().numpy()


def test_mixed_batch_matches_oracle_and_fp64_fallback():
    r = cbkt.rng(4200)
    shapes, ill, outside = batchcase.mixed_mix(r, nuneven=400)
    desc, ap, bp, As, Bs = batchcase.make(r, shapes)
    for k in ill:
        m, n = shapes[k][:2]
        As[k] = batchcase.ill_conditioned(r, m, n)
        batchcase.set_a(ap, desc[k], As[k])
    print(f"case dsgels_batch nprob={len(desc)}", flush=True)
    plan, outs = _mixed(desc, ap, bp, reps=2)
    apg, xpg, itg, infog = outs[0]
    # a replay of the same graph gives the same bits
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)
    apo, xpo, ito, infoo = batchcase.oracle_mixed("N", desc, ap, bp)
    app, bpp, infop = _plain(desc, ap, bp)
    assert np.array_equal(infog, infoo)
    assert np.all(itg[outside] == -1) and np.all(ito[outside] == -1)
    assert np.all(itg[ill] == -31) and np.all(ito[ill] == -31), (itg[ill], ito[ill])
    eps = 2.0 ** -53
    for k, d in enumerate(desc):
        m, n, q = d[:3]
        if itg[k] < 0:
            assert ito[k] < 0 or k not in outside
            # the FP64 batch kernels on X := B, bit for bit
            assert np.array_equal(batchcase.get_a(apg, d), batchcase.get_a(app, d)), (k, d)
            assert np.array_equal(batchcase.get_b(xpg, d), batchcase.get_b(bpp, d)), (k, d)
            assert infog[k] == infop[k]
            continue
        assert ito[k] >= 0, (k, d, itg[k], ito[k])
        assert np.array_equal(batchcase.get_a(apg, d), As[k]), (k, d)
        X = batchcase.get_b(xpg, d)[:n]
        Xo = batchcase.get_b(xpo, d)[:n]
        s = np.linalg.svd(As[k], compute_uv=False)
        kappa = s[0] / s[-1]
        res = np.linalg.norm(Bs[k] - As[k] @ Xo)
        bound = 20 * m * eps * (kappa + kappa ** 2 * res / (s[0] * np.linalg.norm(Xo)))
        assert np.abs(X - Xo).max() <= bound * np.abs(Xo).max(), (k, d)
    assert np.sum(itg >= 0) >= len(desc) - len(ill) - len(outside) - 3


def test_mixed_batch_k5_distribution_backward_error():
    r = cbkt.rng(4201)
    shapes = batchcase.uneven_shapes(r, 3000)
    desc, ap, bp, As, Bs = batchcase.make(r, shapes)
    print(f"case dsgels_batch K5 distribution nprob={len(desc)}", flush=True)
    _, outs = _mixed(desc, ap, bp)
    apg, xpg, itg, infog = outs[0]
    assert np.all(infog == 0)
    worst, at = 0.0, None
    for k, d in enumerate(desc):
        m, n, q = d[:3]
        A, B = As[k], Bs[k]
        X = batchcase.get_b(xpg, d)[:n]
        R = B - A @ X
        an = np.linalg.norm(A)
        for j in range(q):
            eta = np.linalg.norm(A.T @ R[:, j]) / (an * (an * np.linalg.norm(X[:, j]) +
                                                         np.linalg.norm(B[:, j])))
            if eta > worst:
                worst, at = eta, (k, d, eta, int(itg[k]))
    assert worst <= 1e-15, at
    inside = np.array([d[1] <= 32 for d in desc])
    assert np.all(itg[~inside] == -1)
    assert np.mean(itg[inside] >= 0) > 0.99, np.unique(itg, return_counts=True)


def test_mixed_batch_new_buffers_rerecord_and_transposed():
    torch = _t()
    r = cbkt.rng(4202)
    shapes = [(16, 8, 2), (40, 31, 5), (64, 20, 8), (70, 10, 2)]
    desc, ap, bp, As, Bs = batchcase.make(r, shapes)
    print("case dsgels_batch re-record", flush=True)
    plan = cbk_amd.GelsBatchPlan("M", "N", desc)
    work = torch.empty(max(plan.worksize, 1), dtype=torch.uint8, device="cuda")
    res = []
    for _ in range(2):  # fresh A / B / X / iter / info buffers each time
        dA = torch.from_numpy(ap).cuda()
        dB = torch.from_numpy(bp).cuda()
        dX = torch.full_like(dB, 7.0)
        it = torch.zeros(len(desc), dtype=torch.int32, device="cuda")
        info = torch.zeros(len(desc), dtype=torch.int32, device="cuda")
        assert plan.run_mixed(dA, dB, dX, it, info, work) == 0
        assert cbk_amd.synchronize() == 0
        res.append((dA.cpu().numpy(), dX.cpu().numpy(), it.cpu().numpy()))
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)
    assert res[0][2][3] == -1 and np.all(res[0][2][:3] >= 0)
    # transposed: every problem takes the FP64 path, as dgels_batch
    _, outs = _mixed(desc, ap, bp, trans="T")
    apg, xpg, itg, _ = outs[0]
    app, bpp, _ = _plain(desc, ap, bp, trans="T")
    assert np.all(itg == -1)
    assert np.array_equal(apg, app)
    for d in desc:
        assert np.array_equal(batchcase.get_b(xpg, d), batchcase.get_b(bpp, d))


def test_mixed_batch_argument_codes():
    torch = _t()
    r = cbkt.rng(4203)
    desc, ap, bp, _, _ = batchcase.make(r, [(8, 4, 1)])
    plan = cbk_amd.GelsBatchPlan("M", "N", desc)
    plain = cbk_amd.GelsBatchPlan("D", "N", desc)
    dA = torch.from_numpy(ap).cuda()
    dB = torch.from_numpy(bp).cuda()
    dX = torch.empty_like(dB)
    it = torch.zeros(1, dtype=torch.int32, device="cuda")
    w = torch.empty(max(plan.worksize, 1), dtype=torch.uint8, device="cuda")
    f = cbk_amd._bind_batch().dsgels_batch
    P = cbk_amd._ptr
    assert f(plain._h, P(dA), P(dB), P(dX), P(it), P(it), P(w), plan.worksize) == -1
    assert f(plan._h, None, P(dB), P(dX), P(it), P(it), P(w), plan.worksize) == -2
    assert f(plan._h, P(dA), None, P(dX), P(it), P(it), P(w), plan.worksize) == -3
    assert f(plan._h, P(dA), P(dB), None, P(it), P(it), P(w), plan.worksize) == -4
    assert f(plan._h, P(dA), P(dB), P(dX), None, P(it), P(w), plan.worksize) == -5
    assert f(plan._h, P(dA), P(dB), P(dX), P(it), None, P(w), plan.worksize) == -6
    assert f(plan._h, P(dA), P(dB), P(dX), P(it), P(it), None, plan.worksize) == -7
    assert f(plan._h, P(dA), P(dB), P(dX), P(it), P(it), P(w), plan.worksize - 1) == -8
    # the plain entry points refuse a mixed plan
    assert plan.run(dA, dB, it, w) == -1
