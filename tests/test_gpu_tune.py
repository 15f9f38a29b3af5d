"""The kernel-family tuner of the compact routines (cbk_autotune_compact,
include/cbk_ext.h) on the GPU: it times the families that can serve a shape,
installs one that can, reports the V it chose when asked to pick one; and
every family forced through the table (cbk_set_kernel_family) gives the
routine's results against the oracle -- factors bit for bit where the
families share the reference's operation order, solutions within the
routine's rounding bound. The table is cleared after each test."""

import numpy as np
import pytest

import cases
import cbk_amd
import cbkt
import gpuutil

pytestmark = pytest.mark.gpu

F = cbk_amd.FAMILY

SHAPES = [
    # prec, routine, m, n, nrhs, uplo, trans, the case's routine / options / dims
    ("d", "posv", 24, 24, 4, "L", "N", "posv", {"layout": "C", "uplo": "L"}, {"n": 24, "nrhs": 4}),
    ("s", "posv", 16, 16, 2, "U", "N", "posv", {"layout": "C", "uplo": "U"}, {"n": 16, "nrhs": 2}),
    ("d", "sysvnp", 32, 32, 4, "L", "N", "sysvnp", {"layout": "C", "uplo": "L"}, {"n": 32, "nrhs": 4}),
    ("d", "trsm", 24, 12, 0, "L", "N", "trsm",
     {"layout": "C", "side": "L", "uplo": "L", "trans": "N", "diag": "N"}, {"m": 24, "n": 12}),
    ("s", "trsm", 16, 8, 0, "U", "N", "trsm",
     {"layout": "C", "side": "L", "uplo": "U", "trans": "N", "diag": "N"}, {"m": 16, "n": 8}),
    ("d", "gels", 16, 8, 2, "L", "N", "gels", {"layout": "C", "trans": "N"}, {"m": 16, "n": 8, "nrhs": 2}),
    ("d", "gels", 64, 32, 2, "L", "N", "gels", {"layout": "C", "trans": "N"}, {"m": 64, "n": 32, "nrhs": 2}),
    ("d", "geqrf", 32, 10, 0, "L", "N", "geqrf", {"layout": "C"}, {"m": 32, "n": 10}),
    ("d", "geqrf", 64, 32, 0, "L", "N", "geqrf", {"layout": "C"}, {"m": 64, "n": 32}),
]


def _tol(p):
    return 1e-10 if p == "d" else 2e-4


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}{s[1]}_{s[2]}x{s[3]}_{s[4]}")
def test_tuner_installs_a_family_and_every_family_is_right(shape):
    p, routine, m, n, nrhs, uplo, trans, crout, opts, dims = shape
    V = 64
    print(f"case tune {p}{routine} {m}x{n} nrhs={nrhs} {uplo}{trans}", flush=True)
    gpuutil.torch()
    try:
        st, fam, v = cbk_amd.autotune_compact(p, routine, m, n, nrhs, V, nm=16384, uplo=uplo,
                                              trans=trans)
        assert st == 0 and v == V and F["register"] <= fam <= F["grid"], (st, fam, v)
        assert cbk_amd.get_kernel_family(p, routine, m, n, nrhs, V, uplo=uplo, trans=trans) == fam
        print("tuned family", fam, flush=True)
        # every family forced: the oracle's results
        nm = 3 * V + 5
        case = cases.make_case(crout, p, opts, dims, V, nm, seed=77)
        so, outs_o = cases.run(cbkt.oracle().call, case)
        assert so == 0
        for f in (F["register"], F["quad"], F["team"], F["columns"], F["wave"], F["lane"],
                  F["grid"]):
            assert cbk_amd.set_kernel_family(p, routine, m, n, nrhs, V, f, uplo=uplo,
                                             trans=trans) == 0
            sg, outs_g = cases.run(gpuutil.gpu_call, case)
            assert sg == 0
            for o, g in zip(outs_o, outs_g):
                scale = max(1.0, float(np.abs(o).max()))
                assert np.abs(o - g).max() <= _tol(p) * scale, (f, np.abs(o - g).max())
    finally:
        cbk_amd.set_kernel_family(p, routine, m, n, nrhs, V, F["default"], uplo=uplo, trans=trans)


def test_tuner_picks_v():
    gpuutil.torch()
    try:
        st, fam, v = cbk_amd.autotune_compact("d", "posv", 8, 8, 1, 0, nm=32768)
        assert st == 0 and v in (16, 32, 64) and fam >= F["register"], (st, fam, v)
        for vv in (16, 32, 64):
            assert cbk_amd.get_kernel_family("d", "posv", 8, 8, 1, vv) >= F["register"]
        print("tuned V", v, "family", fam, flush=True)
    finally:
        for vv in (16, 32, 64):
            cbk_amd.set_kernel_family("d", "posv", 8, 8, 1, vv, F["default"])
