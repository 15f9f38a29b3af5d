"""The MKL-style adapter (include/cbk_compat.h) on the GPU: every cbk_?..._compact
routine against the include/cbk.h routine it forwards to, bit for bit, on the
same inputs, with the interleave width V taken from the MKL format (SSE / AVX /
AVX512: 16 / 32 / 64 bytes over the scalar size; CBK_COMPACT_V64: V = 64); the
MKL enum decoding (row/column major, upper/lower, left/right, transposes,
unit diagonal); the workspace queries (lwork = -1 writes 1, and gels the size
of its tau scratch, to work[0] on the device); info = -1 for an unrecognized
format with nothing written; the quick returns the reference's adapter keeps
(reference src/cbk_compat.cpp:58-250). The cbk.h routines themselves are
checked against the oracle in the other GPU tests."""

import numpy as np
import pytest

import cases
import cbkt

pytestmark = pytest.mark.gpu

MKL = {"ROW_MAJOR": 101, "COL_MAJOR": 102, "NOTRANS": 111, "TRANS": 112, "CONJTRANS": 113,
       "UPPER": 121, "LOWER": 122, "NONUNIT": 131, "UNIT": 132, "LEFT": 141, "RIGHT": 142,
       "SSE": 181, "AVX": 182, "AVX512": 183, "V64": 190}


def fmt_for(p, V):
    size = 8 if p == "d" else 4
    return {16 // size: MKL["SSE"], 32 // size: MKL["AVX"], 64 // size: MKL["AVX512"]}.get(
        V, MKL["V64"] if V == 64 else None)


def compat_args(case, fmt, work):
    """the cbk_compat.h argument list of a cbk.h case (None in the info slot)"""
    a = case["args"]
    r = case["routine"]
    L = MKL["ROW_MAJOR"] if a[0] in "Rr" else MKL["COL_MAJOR"]
    if r == "geqrf":
        layout, m, n, ap, lda, taup, V, nm = a
        return [L, m, n, ap, lda, taup, work, 1, None, fmt, nm]
    if r == "ormqr":
        layout, side, tr, m, n, k, ap, lda, taup, cp, ldc, V, nm = a
        tr = case["opts"].get("compat_trans", tr)  # 'C' reaches the adapter as a transpose
        return [L, side, tr, m, n, k, ap, lda, taup, cp, ldc, work, 1, None, fmt, nm]
    if r == "orgqr":
        layout, m, n, k, ap, lda, taup, V, nm = a
        return [L, m, n, k, ap, lda, taup, work, 1, None, fmt, nm]
    if r in ("potrf", "sytrfnp"):
        layout, uplo, n, ap, lda, V, nm = a
        U = MKL["UPPER"] if uplo in "Uu" else MKL["LOWER"]
        return [L, U, n, ap, lda, None, fmt, nm]
    if r in ("potrs", "posv", "sytrsnp", "sysvnp"):
        layout, uplo, n, nrhs, ap, lda, bp, ldb, V, nm = a
        U = MKL["UPPER"] if uplo in "Uu" else MKL["LOWER"]
        return [L, U, n, nrhs, ap, lda, bp, ldb, None, fmt, nm]
    if r == "trsm":
        layout, side, uplo, tr, diag, m, n, alpha, ap, lda, bp, ldb, V, nm = a
        return [L, MKL["LEFT"] if side in "Ll" else MKL["RIGHT"],
                MKL["UPPER"] if uplo in "Uu" else MKL["LOWER"],
                MKL["NOTRANS"] if tr in "Nn" else (MKL["TRANS"] if tr in "Tt" else MKL["CONJTRANS"]),
                MKL["UNIT"] if diag in "Uu" else MKL["NONUNIT"], m, n, alpha, ap, lda, bp, ldb,
                fmt, nm]
    if r == "gels":
        layout, tr, m, n, nrhs, ap, lda, bp, ldb, taup, V, nm = a
        return [L, tr, m, n, nrhs, ap, lda, bp, ldb, taup, len(taup), None, fmt, nm]
    raise ValueError(r)


def run_pair(case):
    """(outputs of the cbk.h call, outputs of the compat call, compat info)"""
    import cbk_amd
    import gpuutil
    p = case["p"]
    fmt = fmt_for(p, case["V"])
    # the cbk.h routine
    args = [gpuutil.to_dev(x) if isinstance(x, np.ndarray) else x for x in case["args"]]
    st = cbk_amd.call(p, case["routine"], *args)
    assert st == 0
    ref = [args[i].cpu().numpy() for i in case["outs"]]
    # the adapter, on fresh copies
    work = gpuutil.to_dev(np.full(4, -7.0, dtype=cbkt.DT[p]))
    dev = {}
    cargs = compat_args(case, fmt, work)
    for j, x in enumerate(case["args"]):
        if isinstance(x, np.ndarray):
            dev[j] = gpuutil.to_dev(x)
    # substitute the device copies for the numpy buffers, in order
    it = iter(dev[j] for j in sorted(dev))
    cargs = [next(it) if isinstance(x, np.ndarray) else x for x in cargs]
    info = cbk_amd.call_compat(p, case["routine"], *cargs)
    assert cbk_amd.synchronize() == 0
    got = [dev[i].cpu().numpy() for i in case["outs"]]
    return ref, got, info


CASES = [
    ("geqrf", {"layout": "C"}, {"m": 12, "n": 5}),
    ("geqrf", {"layout": "R"}, {"m": 6, "n": 9}),
    ("ormqr", {"layout": "C", "side": "L", "trans": "T"}, {"m": 10, "n": 3, "k": 4}),
    ("ormqr", {"layout": "R", "side": "R", "trans": "T", "compat_trans": "C"}, {"m": 3, "n": 8, "k": 5}),
    ("orgqr", {"layout": "C"}, {"m": 9, "n": 6, "k": 4}),
    ("orgqr", {"layout": "C"}, {"m": 5, "n": 3, "k": 0}),
    ("potrf", {"layout": "C", "uplo": "U"}, {"n": 13}),
    ("potrs", {"layout": "R", "uplo": "L"}, {"n": 9, "nrhs": 3}),
    ("posv", {"layout": "C", "uplo": "L"}, {"n": 32, "nrhs": 4}),
    ("posv", {"layout": "C", "uplo": "U"}, {"n": 7, "nrhs": 0}),
    ("sytrfnp", {"layout": "R", "uplo": "U"}, {"n": 10}),
    ("sytrsnp", {"layout": "C", "uplo": "L"}, {"n": 8, "nrhs": 2}),
    ("sysvnp", {"layout": "C", "uplo": "L"}, {"n": 24, "nrhs": 5}),
    ("trsm", {"layout": "C", "side": "L", "uplo": "L", "trans": "N", "diag": "U"},
     {"m": 12, "n": 5, "alpha": 0.75}),
    ("trsm", {"layout": "R", "side": "R", "uplo": "U", "trans": "C", "diag": "N"},
     {"m": 4, "n": 9, "alpha": -1.5}),
    ("trsm", {"layout": "C", "side": "L", "uplo": "U", "trans": "T", "diag": "N"},
     {"m": 6, "n": 2, "alpha": 0.0}),
    ("gels", {"layout": "C", "trans": "N"}, {"m": 16, "n": 8, "nrhs": 2}),
    ("gels", {"layout": "R", "trans": "C"}, {"m": 7, "n": 11, "nrhs": 3}),
]


@pytest.mark.parametrize("p", ["d", "s"])
@pytest.mark.parametrize("fmt_name", ["SSE", "AVX", "AVX512", "V64"])
@pytest.mark.parametrize("routine,opts,dims", CASES)
def test_compat_matches_cbk(routine, opts, dims, p, fmt_name):
    size = 8 if p == "d" else 4
    V = 64 if fmt_name == "V64" else {"SSE": 16, "AVX": 32, "AVX512": 64}[fmt_name] // size
    nm = 2 * V + 3
    print(f"case compat {p}{routine} {opts} {dims} {fmt_name} V={V}", flush=True)
    seed = 31 * sum(map(ord, routine)) + 7 * V + (p == "d") + sum(map(ord, str(sorted(opts.items()))))
    case = cases.make_case(routine, p, opts, dims, V, nm, seed=seed % 100000)
    ref, got, info = run_pair(case)
    assert info in (0, None)
    for r, g in zip(ref, got):
        assert cases.same_bits(r, g), (routine, p, fmt_name)


@pytest.mark.parametrize("p", ["d", "s"])
def test_compat_queries_and_bad_format(p):
    import cbk_amd
    import gpuutil
    dt = cbkt.DT[p]
    work = gpuutil.to_dev(np.full(2, -7.0, dtype=dt))
    z = 0
    # geqrf / ormqr / orgqr: lwork = -1 writes 1
    assert cbk_amd.call_compat(p, "geqrf", MKL["COL_MAJOR"], 8, 4, z, 8, z, work, -1, None,
                               MKL["AVX"], 10) == 0
    assert cbk_amd.synchronize() == 0 and work.cpu().numpy()[0] == 1.0
    work.fill_(-7.0)
    assert cbk_amd.call_compat(p, "ormqr", MKL["COL_MAJOR"], "L", "T", 8, 2, 4, z, 8, z, z, 8,
                               work, -1, None, MKL["SSE"], 10) == 0
    assert cbk_amd.synchronize() == 0 and work.cpu().numpy()[0] == 1.0
    work.fill_(-7.0)
    assert cbk_amd.call_compat(p, "orgqr", MKL["ROW_MAJOR"], 8, 4, 4, z, 4, z, work, -1, None,
                               MKL["V64"], 10) == 0
    assert cbk_amd.synchronize() == 0 and work.cpu().numpy()[0] == 1.0
    # gels: the tau scratch, min(m,n) * V * ceil(nm / V)
    size = 8 if p == "d" else 4
    for fmt, V in ((MKL["SSE"], 16 // size), (MKL["AVX512"], 64 // size), (MKL["V64"], 64)):
        work.fill_(-7.0)
        assert cbk_amd.call_compat(p, "gels", MKL["COL_MAJOR"], "N", 12, 5, 2, z, 12, z, 12,
                                   work, -1, None, fmt, 37) == 0
        assert cbk_amd.synchronize() == 0
        assert work.cpu().numpy()[0] == 5 * V * ((37 + V - 1) // V)
    assert cbk_amd.lib().cbk_compact_vlen(MKL["AVX"], 1 if p == "d" else 0) == 32 // size
    assert cbk_amd.lib().cbk_compact_vlen(999, 1) == 0
    # an unrecognized format: info -1, nothing written
    V = 4
    case = cases.make_case("posv", p, {"layout": "C", "uplo": "L"}, {"n": 6, "nrhs": 2}, V, 9,
                           seed=5)
    a = gpuutil.to_dev(case["args"][4])
    b = gpuutil.to_dev(case["args"][6])
    assert cbk_amd.call_compat(p, "posv", MKL["COL_MAJOR"], MKL["LOWER"], 6, 2, a, 6, b, 6,
                               None, 175, 9) == -1
    assert cbk_amd.synchronize() == 0
    assert np.array_equal(a.cpu().numpy(), case["args"][4])
    assert np.array_equal(b.cpu().numpy(), case["args"][6])
    # an inconsistent leading dimension: the cbk.h code of the forwarded call
    assert cbk_amd.call_compat(p, "posv", MKL["COL_MAJOR"], MKL["LOWER"], 6, 2, a, 5, b, 6,
                               None, MKL["AVX"], 9) == -6
    assert np.array_equal(a.cpu().numpy(), case["args"][4])
