"""GPU edge contracts against the oracle, bit for bit.

* trsm with alpha == 0: B := 0 in every group without referencing A
  (reference src/cbk_trsm_compact.hpp:286-289) -- A is passed as NULL to the
  library and to the oracle alike; B's leading-dimension padding and the
  padded group slots' rows beyond m stay as they were; alpha = -0.0 is zero
  too and writes +0.0.
* A leading dimension too large for the register and team kernels' 32-bit
  buffer offsets (the dispatcher's offsets_fit): the call must take the lane
  kernels and still give the oracle's bits. The operand is a 2.5 GB device
  buffer of which only the referenced elements are set.
"""

import numpy as np
import pytest

import cases
import cbkt

pytestmark = pytest.mark.gpu


def _gpu():
    import gpuutil
    return gpuutil


@pytest.mark.parametrize("p", ["d", "s"])
@pytest.mark.parametrize("V,nm", [(64, 130), (32, 71), (2, 5)])
def test_trsm_alpha_zero_never_reads_a(p, V, nm):
    g = _gpu()
    for i, (lay, side, uplo, tr, diag) in enumerate([("C", "L", "L", "N", "U"), ("R", "R", "U", "T", "N"),
                                                     ("C", "R", "L", "C", "N"), ("R", "L", "U", "N", "U")]):
        for m, n, alpha in ((24, 24, 0.0), (5, 3, -0.0), (16, 1, 0.0)):
            print(f"case trsm{p} alpha={alpha} {lay}{side}{uplo}{tr}{diag} m={m} n={n} V={V} nm={nm}",
                  flush=True)
            c = cases.make_case("trsm", p, {"layout": lay, "side": side, "uplo": uplo, "trans": tr,
                                            "diag": diag}, {"m": m, "n": n, "alpha": alpha}, V, nm,
                                seed=12000 + i + m, ld_extra=1)
            c["args"][8] = None  # A: never referenced
            b0 = c["args"][10].copy()
            so, (bo,) = cases.run(cbkt.oracle().call, c)
            sg, (bg,) = cases.run(g.gpu_call, c)
            assert so == sg == 0
            assert cases.same_bits(bo, bg)
            row = lay == "R"
            ldb = c["args"][11]
            Bg = cbkt.unpack(bg, nm, m, n, V, ld=ldb, row=row, all_slots=True)
            assert np.all(cases.bits(Bg) == 0), "B must be +0.0 everywhere"
            assert np.array_equal(bg == cbkt.SENTINEL, b0 == cbkt.SENTINEL)


def _large_ld_case(routine, ld_big):
    """posv (or dsposv) n = 16, V = 64, one group, with ldap = ld_big"""
    import torch
    n, nrhs, V, nm = 16, 4, 64, 64
    c = cases.make_case(routine, "d", {"layout": "C", "uplo": "L"}, {"n": n, "nrhs": nrhs}, V, nm,
                        seed=123)
    small = c["args"][4]  # ldap = n
    i, j, v = np.meshgrid(np.arange(n), np.arange(n), np.arange(V), indexing="ij")
    idx_small = ((i + j * n) * V + v).reshape(-1)
    idx_big = ((i + j * ld_big) * V + v).reshape(-1)
    big = torch.zeros(ld_big * n * V, dtype=torch.float64, device="cuda")
    big[torch.from_numpy(idx_big).cuda()] = torch.from_numpy(small[idx_small]).cuda()
    return c, big, torch.from_numpy(idx_small), torch.from_numpy(idx_big).cuda()


def test_posv_large_leading_dimension_takes_64bit_path():
    import cbk_amd
    g = _gpu()
    torch = g.torch()
    ld_big = 300000  # (15 + 15 * ld) * 64 * 8 bytes > 2^31
    assert (15 + 15 * ld_big) * 64 * 8 > 2 ** 31
    c, big, idx_small, idx_big = _large_ld_case("posv", ld_big)
    print("case posv n=16 V=64 ldap=300000", flush=True)
    so, (ao, bo) = cases.run(cbkt.oracle().call, c)
    assert so == 0
    args = list(c["args"])
    db = g.to_dev(args[6])
    st = cbk_amd.call("d", "posv", "C", "L", 16, 4, big, ld_big, db, 16, 64, 64)
    assert st == 0 and cbk_amd.synchronize() == 0
    a_small = ao.copy()
    a_small[idx_small.numpy()] = big[idx_big].cpu().numpy()
    assert cases.same_bits(a_small, ao)
    assert cases.same_bits(db.cpu().numpy(), bo)
    del big
    torch.cuda.empty_cache()
