"""GPU parity of the register lane kernel (potrf/posv/potrs, sytrfnp/sysvnp/
sytrsnp at n <= 16 in FP64, n <= 24 in FP32) against the oracle, bit for bit
-- factors and solutions:
the kernel performs the lane kernel's operations in registers, including the
reference trsm router's choice of the axpy form for a single leftover
right-hand side of the backward sweep (nrhs = 1, 3, 5 with column-major B).
Also the per-matrix info codes with a failed matrix keeping its B."""

import numpy as np
import pytest

import cases
import cbkt

pytestmark = pytest.mark.gpu


def _gpu():
    import gpuutil
    return gpuutil


class _Register:
    """route the shapes the quad kernel serves by default (FP32 order 24) to
    the register kernel through the kernel-family table, for the duration"""

    def __init__(self, p, n, nrhs_list, V_list):
        self.keys = [(p, r, n, n, nrhs, V) for r in ("posv", "sysvnp") for nrhs in nrhs_list
                     for V in V_list] if (p, n) == ("s", 24) else []

    def __enter__(self):
        import cbk_amd
        for k in self.keys:
            assert cbk_amd.set_kernel_family(*k, cbk_amd.FAMILY["register"]) == 0

    def __exit__(self, *exc):
        import cbk_amd
        for k in self.keys:
            cbk_amd.set_kernel_family(*k, cbk_amd.FAMILY["default"])


@pytest.mark.parametrize("p,n", [(p, n) for p in "ds" for n in (1, 3, 8, 9, 12, 13, 16)] +
                         [("s", 20), ("s", 24)])
def test_reg_factor_and_solve_bitwise(p, n):
    with _Register(p, n, (0, 1, 3, 4, 5), (64, 32, 8, 2)):
        _reg_factor_and_solve_bitwise(p, n)


def _reg_factor_and_solve_bitwise(p, n):
    for i, (lay, uplo) in enumerate((("C", "L"), ("C", "U"), ("R", "L"), ("R", "U"))):
        for V, nm in ((64, 130), (32, 71), (8, 13), (2, 5)):
            for rt in ("potrf", "sytrfnp", "posv", "sysvnp", "potrs", "sytrsnp"):
                for nrhs in ((0,) if rt in ("potrf", "sytrfnp") else (1, 3, 4, 5)):
                    dims = {"n": n} if nrhs == 0 else {"n": n, "nrhs": nrhs}
                    print(f"case reg {rt}{p} {lay}{uplo} n={n} nrhs={nrhs} V={V} nm={nm}", flush=True)
                    c = cases.make_case(rt, p, {"layout": lay, "uplo": uplo}, dims, V, nm,
                                        seed=9100 + 7 * n + i + nrhs, ld_extra=i % 2)
                    so, outo = cases.run(cbkt.oracle().call, c)
                    sg, outg = cases.run(_gpu().gpu_call, c)
                    assert so == sg == 0
                    for a, b in zip(outo, outg):
                        assert cases.same_bits(a, b), (rt, lay, uplo, V, nrhs)


@pytest.mark.parametrize("p,n", [("d", 8), ("d", 16), ("s", 24), ("d", 13)])
def test_reg_info(p, n):
    """info = k + 1 for a failed pivot k (Cholesky: a negative diagonal; LDL^T:
    row and column k zeroed up to the diagonal, an exact zero D(k)), every k
    of the order over the batch; failed matrices keep B"""
    with _Register(p, n, (2,), (64,)):
        _reg_info(p, n)


def _reg_info(p, n):
    V, nm = 64, 70
    for rt in ("posv", "sysvnp"):
        c = cases.make_case(rt, p, {"layout": "C", "uplo": "L"}, {"n": n, "nrhs": 2}, V, nm,
                            seed=4 + n)
        A = cbkt.unpack(c["args"][4], nm, n, n, V)
        bad = {3 + 2 * k: k for k in range(n)}
        for mat, k in bad.items():
            if rt == "posv":
                A[mat, k, k] = -3.0
            else:
                A[mat, k, :k + 1] = 0.0
                A[mat, :k + 1, k] = 0.0
        c["args"][4] = cbkt.pack(A, V)
        info_o = np.zeros(nm, np.int32)
        info_g = np.zeros(nm, np.int32)
        print(f"case reg info {rt}{p} n={n}", flush=True)
        so, outo = cases.run(cbkt.oracle().call, c, info=info_o)
        sg, outg = cases.run(_gpu().gpu_call, c, info=info_g)
        assert so == sg == 0
        assert np.array_equal(info_o, info_g)
        want = np.zeros(nm, np.int32)
        for mat, k in bad.items():
            want[mat] = k + 1
        assert np.array_equal(info_g, want), (info_g, want)
        for a, b in zip(outo, outg):
            assert cases.same_bits_nan(a, b)
