"""GPU parity of the uneven-batch gels (dgels_batch / sgels_batch) against the
oracle's batch gels on the same pools.

Every kernel class is exercised: the team kernel (n + nrhs <= 16; rows <= 16,
32, 64), the wavefront classes (n + nrhs <= 64; rows <= 32, 64: the lane-grid
kernel for n <= 32 and nrhs <= 8, two problems per wavefront; the 8 x 8 lane
grid for wider problems with nrhs <= 8, one per wavefront; otherwise two
problems per wavefront when n + nrhs <= 32, one otherwise), the
two-array wavefront kernel (n + nrhs > 64; with nrhs <= 8 the 8 x 8 lane grid)
and the general lane kernel
(transposed, underdetermined, more than 64 rows, empty problems). The general kernel performs the oracle's
operations and is compared bit for bit; the register kernels sum in another
order and are held to the least-squares perturbation bound (batchcase.check_ls).
Also: info codes, untouched padding, replays of the recorded graph, a
re-recorded graph for new buffers, the tuning switch between kernels."""

import numpy as np
import pytest

import batchcase
import cases
import cbk_amd
import cbkt

pytestmark = pytest.mark.gpu


def _t():
    import gpuutil
    return gpuutil.torch()


def _run(p, trans, desc, ap, bp, with_info=True):
    torch = _t()
    plan = cbk_amd.GelsBatchPlan("D" if p == "d" else "S", trans, desc)
    dA = torch.from_numpy(ap).cuda()
    dB = torch.from_numpy(bp).cuda()
    info = torch.full((len(desc),), -9, dtype=torch.int32, device="cuda") if with_info else None
    work = torch.empty(max(plan.worksize, 1), dtype=torch.uint8, device="cuda")
    st = plan.run(dA, dB, info, work)
    assert st == 0, (st, cbk_amd.last_error())
    assert cbk_amd.synchronize() == 0, cbk_amd.last_error()
    out = (dA.cpu().numpy(), dB.cpu().numpy(), info.cpu().numpy() if with_info else None)
    return plan, out


def _class_mix(r):
    return [
        (16, 8, 2), (12, 5, 3), (16, 15, 1), (4, 2, 1),          # team, 16 rows
        (32, 12, 4), (20, 10, 6), (31, 14, 2),                   # team, 32 rows
        (64, 8, 8), (50, 13, 3), (64, 15, 1),                    # team, 64 rows
        (32, 30, 2), (24, 20, 8), (32, 32, 8),                   # wavefront, 32 rows
        (64, 32, 2), (64, 56, 8), (40, 33, 5), (63, 60, 4),      # wavefront, 64 rows
        (64, 20, 4), (50, 25, 7), (33, 17, 1), (30, 9, 8),       # two per wavefront (n + nrhs <= 32)
        (64, 64, 1), (64, 60, 8), (60, 58, 7), (64, 57, 8),      # two-array wavefront
        (30, 12, 16), (32, 20, 10), (50, 20, 9),                 # nrhs > 8: two per wavefront
        (60, 30, 12), (40, 16, 20), (32, 20, 14),                # nrhs > 8: one per wavefront
        (64, 56, 10), (60, 50, 15),                              # nrhs > 8: two-array wavefront
        (70, 10, 2), (30, 40, 2), (5, 0, 2), (6, 3, 0),          # general
    ] + batchcase.uneven_shapes(r, 300)


@pytest.mark.parametrize("p", ["d", "s"])
def test_batch_all_classes(p):
    r = cbkt.rng(900 + (p == "s"))
    shapes = _class_mix(r)
    print(f"case batch {p} nprob={len(shapes)}", flush=True)
    desc, ap, bp, As, Bs = batchcase.make(r, shapes, cbkt.DT[p])
    apo, bpo = ap.copy(), bp.copy()
    info_o = np.zeros(len(desc), np.int32)
    assert batchcase.oracle_batch(p, "N", desc, apo, bpo, info_o) == 0
    plan, (apg, bpg, info_g) = _run(p, "N", desc, ap, bp)
    cls = plan.classes()
    assert all(c > 0 for c in cls), cls
    assert np.array_equal(info_g, info_o)
    batchcase.check_ls(p, desc, As, Bs, apo, bpo, apg, bpg)
    # general-kernel problems: the oracle's operations exactly
    for k, d in enumerate(desc):
        m, n, q = d[:3]
        if m > 64 or m < n or n == 0 or q == 0:
            assert cases.same_bits(batchcase.get_a(apg, d), batchcase.get_a(apo, d)), d
            assert cases.same_bits(batchcase.get_b(bpg, d), batchcase.get_b(bpo, d)), d
    # nothing outside the problems is written
    assert np.array_equal(apg == 7.0, apo == 7.0) and np.array_equal(bpg == 7.0, bpo == 7.0)


@pytest.mark.parametrize("trans", ["T", "C"])
def test_batch_transposed_is_general_and_exact(trans):
    r = cbkt.rng(910)
    shapes = batchcase.uneven_shapes(r, 80, mmax=24, rmax=3) + [(9, 4, 2), (4, 9, 1)]
    print(f"case batch trans={trans} nprob={len(shapes)}", flush=True)
    desc, ap, bp, As, Bs = batchcase.make(r, shapes)
    apo, bpo = ap.copy(), bp.copy()
    info_o = np.zeros(len(desc), np.int32)
    assert batchcase.oracle_batch("d", trans, desc, apo, bpo, info_o) == 0
    plan, (apg, bpg, info_g) = _run("d", trans, desc, ap, bp)
    assert plan.classes()[6] == len(desc)
    assert cases.same_bits(apg, apo) and cases.same_bits(bpg, bpo)
    assert np.array_equal(info_g, info_o)


def test_batch_info_zero_columns():
    r = cbkt.rng(920)
    shapes = [(16, 8, 2), (32, 10, 2), (64, 12, 3), (30, 25, 2), (60, 40, 4), (20, 6, 1)]
    desc, ap, bp, As, Bs = batchcase.make(r, shapes)
    for k, col in ((0, 3), (1, 0), (2, 11), (3, 7), (4, 39)):
        m, n, q, lda, ldb, a0, b0 = desc[k]
        ap[a0 + col * lda: a0 + col * lda + m] = 0.0
    apo, bpo = ap.copy(), bp.copy()
    info_o = np.zeros(len(desc), np.int32)
    assert batchcase.oracle_batch("d", "N", desc, apo, bpo, info_o) == 0
    assert info_o.tolist() == [4, 1, 12, 8, 40, 0]
    _, (apg, bpg, info_g) = _run("d", "N", desc, ap, bp)
    assert np.array_equal(info_g, info_o)
    for k in range(5):  # B holds Q^T B: same column norms as B
        d = desc[k]
        assert np.allclose(np.linalg.norm(batchcase.get_b(bpg, d), axis=0),
                           np.linalg.norm(Bs[k], axis=0), rtol=1e-12)


def test_batch_graph_replay_and_new_buffers():
    """the same plan: replays on refreshed inputs give the first results bit for
    bit; new buffers re-record the graph; info may be omitted"""
    torch = _t()
    r = cbkt.rng(930)
    desc, ap, bp, As, Bs = batchcase.make(r, _class_mix(r))
    plan = cbk_amd.GelsBatchPlan("D", "N", desc)
    work = torch.empty(plan.worksize, dtype=torch.uint8, device="cuda")
    A0 = torch.from_numpy(ap).cuda()
    B0 = torch.from_numpy(bp).cuda()
    dA, dB = A0.clone(), B0.clone()
    info = torch.zeros(len(desc), dtype=torch.int32, device="cuda")
    outs = []
    for rep in range(3):
        dA.copy_(A0)
        dB.copy_(B0)
        assert plan.run(dA, dB, info, work) == 0
        torch.cuda.synchronize()
        outs.append((dA.cpu().numpy(), dB.cpu().numpy()))
    for a, b in outs[1:]:
        assert cases.same_bits(a, outs[0][0]) and cases.same_bits(b, outs[0][1])
    dA2, dB2 = A0.clone(), B0.clone()
    assert plan.run(dA2, dB2, None, work) == 0
    torch.cuda.synchronize()
    assert cases.same_bits(dA2.cpu().numpy(), outs[0][0])
    assert cases.same_bits(dB2.cpu().numpy(), outs[0][1])
    assert plan.run(dA2, dB2, None, None, lwork=plan.worksize) == -5
    small = torch.empty(plan.worksize - 8, dtype=torch.uint8, device="cuda")
    assert plan.run(dA2, dB2, None, small) == -6


def test_batch_tuning_routes_and_autotune():
    """narrow_max_cols = 0 sends every team-class problem to the wavefront
    kernel: the classes change, the results stay within the bound; the tuner
    runs and installs a valid table"""
    r = cbkt.rng(940)
    shapes = [(16, 8, 2), (32, 12, 4), (64, 10, 3), (28, 7, 1)] + batchcase.uneven_shapes(r, 60, mmax=32, rmax=4)
    desc, ap, bp, As, Bs = batchcase.make(r, shapes)
    apo, bpo = ap.copy(), bp.copy()
    assert batchcase.oracle_batch("d", "N", desc, apo, bpo) == 0
    saved = cbk_amd.get_tuning()
    try:
        assert cbk_amd.set_tuning([0, 0, 0], [0] * 7) == 0
        plan, (apg, bpg, _) = _run("d", "N", desc, ap.copy(), bp.copy())
        cls = plan.classes()
        assert cls[0] == cls[1] == cls[2] == 0 and cls[3] + cls[4] > 0
        batchcase.check_ls("d", desc, As, Bs, apo, bpo, apg, bpg)
        assert cbk_amd.set_tuning([17, 0, 0], [0] * 7) != 0
        assert cbk_amd.autotune() == 0
        t = cbk_amd.get_tuning()
        assert all(0 <= x <= 16 for x in t["narrow_max_cols"])
        assert all(0 <= x <= 64 for x in t["blocks_per_cu"])
        print("tuned", t, flush=True)
    finally:
        cbk_amd.set_tuning(saved["narrow_max_cols"], saved["blocks_per_cu"])


def test_batch_plan_argument_codes():
    ok = [(4, 2, 1, 4, 4, 0, 0)]
    with pytest.raises(ValueError, match="returned -1"):
        cbk_amd.GelsBatchPlan("Z", "N", ok)
    with pytest.raises(ValueError, match="returned -2"):
        cbk_amd.GelsBatchPlan("D", "X", ok)
    with pytest.raises(ValueError, match=r"returned -5 \(bad descriptor 1\)"):
        cbk_amd.GelsBatchPlan("D", "N", ok + [(4, 2, 1, 3, 4, 0, 0)])
    with pytest.raises(ValueError, match=r"returned -5 \(bad descriptor 0\)"):
        cbk_amd.GelsBatchPlan("D", "N", [(2, 5, 1, 2, 4, 0, 0)])
    plan = cbk_amd.GelsBatchPlan("S", "N", ok)
    assert plan.run(None, None, None, None) in (-2,)
