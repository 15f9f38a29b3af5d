"""Seeded problem cases shared by the oracle-vs-reference, golden-fixture and
GPU parity tests. A case is a routine, its option characters and sizes, and
the packed input buffers; `run(lib_call, case)` executes it through any
library speaking the cbk.h ABI and returns the output buffers."""

import itertools

import numpy as np

import cbkt


def _dt(p):
    return cbkt.DT[p]


def make_case(routine, p, opts, dims, V, nm, seed, ld_extra=0):
    """Build the inputs of one call. opts: dict of option chars; dims: dict.
    Returns a dict with 'args' (the ABI argument list with numpy buffers) and
    the names of the output buffers."""
    r = cbkt.rng(seed)
    dt = _dt(p)
    layout = opts.get("layout", "C")
    row = layout in "Rr"
    c = {"routine": routine, "p": p, "opts": opts, "dims": dims, "V": V, "nm": nm}
    if routine in ("potrf", "sytrfnp", "posv", "potrs", "sysvnp", "sytrsnp"):
        n = dims["n"]
        nrhs = dims.get("nrhs", 0)
        uplo = opts.get("uplo", "L")
        if routine in ("potrf", "posv", "potrs"):
            A = cbkt.gen_spd(r, nm, n, dt)
        else:
            A = cbkt.gen_sym_ldlt(r, nm, n, dt)
        lda = max(1, n) + ld_extra
        ap = cbkt.pack(A, V, ld=lda, row=row, dtype=dt)
        if routine in ("potrs", "sytrsnp"):
            # factor with the oracle first: the solves take a factor
            f = "potrf" if routine == "potrs" else "sytrfnp"
            assert cbkt.oracle().call(p, f, layout, uplo, n, ap, lda, V, nm) == 0
        if routine in ("potrf", "sytrfnp"):
            c["args"] = [layout, uplo, n, ap, lda, V, nm]
            c["outs"] = [3]
        else:
            B = r.standard_normal((nm, n, nrhs)).astype(dt)
            ldb = (max(1, nrhs) if row else max(1, n)) + ld_extra
            bp = cbkt.pack(B, V, ld=ldb, row=row, pad_diag=0.0, dtype=dt)
            c["args"] = [layout, uplo, n, nrhs, ap, lda, bp, ldb, V, nm]
            c["outs"] = [4, 6]
    elif routine == "trsm":
        side, uplo, tr, diag = opts["side"], opts["uplo"], opts["trans"], opts["diag"]
        m, n = dims["m"], dims["n"]
        alpha = dims.get("alpha", 1.0)
        s = m if side in "Ll" else n
        A = cbkt.gen_tri(r, nm, s, uplo in "Uu", dt)
        lda = max(1, s) + ld_extra
        ap = cbkt.pack(A, V, ld=lda, row=row, dtype=dt)
        B = r.uniform(-1, 1, (nm, m, n)).astype(dt)
        ldb = (max(1, n) if row else max(1, m)) + ld_extra
        bp = cbkt.pack(B, V, ld=ldb, row=row, pad_diag=0.0, dtype=dt)
        c["args"] = [layout, side, uplo, tr, diag, m, n, alpha, ap, lda, bp, ldb, V, nm]
        c["outs"] = [10]
    elif routine == "geqrf":
        m, n = dims["m"], dims["n"]
        A = cbkt.gen_uniform(r, nm, m, n, dt)
        lda = (max(1, n) if row else max(1, m)) + ld_extra
        ap = cbkt.pack(A, V, ld=lda, row=row, dtype=dt)
        k = min(m, n)
        taup = np.full(max(1, ((nm + V - 1) // V) * k * V), cbkt.SENTINEL, dtype=dt)
        c["args"] = [layout, m, n, ap, lda, taup, V, nm]
        c["outs"] = [3, 5]
    elif routine in ("ormqr", "orgqr"):
        # reflectors from the oracle's geqrf of a random nq x k (ormqr) / m x n (orgqr)
        if routine == "ormqr":
            side, tr = opts["side"], opts["trans"]
            m, n, k = dims["m"], dims["n"], dims["k"]
            nq = m if side in "Ll" else n
            qr_m, qr_n = nq, k
        else:
            m, n, k = dims["m"], dims["n"], dims["k"]
            qr_m, qr_n = m, n
        A = cbkt.gen_uniform(r, nm, qr_m, qr_n, dt)
        lda = (max(1, qr_n) if row else max(1, qr_m)) + ld_extra
        ap = cbkt.pack(A, V, ld=lda, row=row, dtype=dt)
        kq = min(qr_m, qr_n)
        taup = np.zeros(max(1, ((nm + V - 1) // V) * kq * V), dtype=dt)
        if qr_m > 0 and qr_n > 0 and nm > 0:
            assert cbkt.oracle().call(p, "geqrf", layout, qr_m, qr_n, ap, lda, taup, V, nm) == 0
        if routine == "orgqr" and k < kq:
            # orgqr's tau has ld = k: re-pack the first k scalars
            t = cbkt.unpack_vec(taup, nm, kq, V)[:, :k] if kq > 0 else np.zeros((nm, 0), dt)
            taup = cbkt.pack_vec(t, V, dtype=dt) if k > 0 else np.zeros(1, dt)
        if routine == "ormqr":
            C = r.uniform(-1, 1, (nm, m, n)).astype(dt)
            ldc = (max(1, n) if row else max(1, m)) + ld_extra
            cp = cbkt.pack(C, V, ld=ldc, row=row, pad_diag=0.0, dtype=dt)
            c["args"] = [layout, side, tr, m, n, k, ap, lda, taup, cp, ldc, V, nm]
            c["outs"] = [9]
        else:
            c["args"] = [layout, m, n, k, ap, lda, taup, V, nm]
            c["outs"] = [4]
    elif routine == "gels":
        tr = opts["trans"]
        m, n, nrhs = dims["m"], dims["n"], dims["nrhs"]
        A = cbkt.gen_uniform(r, nm, m, n, dt)
        lda = (max(1, n) if row else max(1, m)) + ld_extra
        ap = cbkt.pack(A, V, ld=lda, row=row, dtype=dt)
        mx = max(m, n)
        B = r.uniform(-1, 1, (nm, mx, nrhs)).astype(dt)
        ldb = (max(1, nrhs) if row else max(1, mx)) + ld_extra
        bp = cbkt.pack(B, V, ld=ldb, row=row, pad_diag=0.0, dtype=dt)
        q = min(m, n)
        taup = np.full(max(1, ((nm + V - 1) // V) * q * V), cbkt.SENTINEL, dtype=dt)
        c["args"] = [layout, tr, m, n, nrhs, ap, lda, bp, ldb, taup, V, nm]
        c["outs"] = [5, 7, 9]
    else:
        raise ValueError(routine)
    return c


def run(call, case, info=None):
    """Run a case through call(p, routine, *args, info=...) on copies of its
    buffers; returns (status, [output buffers])."""
    args = [a.copy() if isinstance(a, np.ndarray) else a for a in case["args"]]
    st = call(case["p"], case["routine"], *args, info=info)
    return st, [args[i] for i in case["outs"]]


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def same_bits(x, y):
    return np.array_equal(bits(x), bits(y))


def same_bits_nan(x, y):
    """bit equality with every NaN taken as one value (NaN payloads and signs
    differ between x86 and the GPU; a poisoned lane is NaN on both)"""
    x, y = np.asarray(x), np.asarray(y)
    nx, ny = np.isnan(x), np.isnan(y)
    if not np.array_equal(nx, ny):
        return False
    return np.array_equal(bits(np.where(nx, 0, x)), bits(np.where(ny, 0, y)))


# ---------------------------------------------------------------- case sweeps

SMALL = (0, 1, 2, 3, 5)


def sweep(quick=False):
    """Yield (routine, p, opts, dims) over the option and size space, the
    reference tests' small dims {0,1,2,3,5} plus sizes crossing the NB=8 panel
    and the 4/2/1 RHS blocking."""
    sizes = (0, 1, 3, 5, 8, 9, 17) if not quick else (1, 5, 9)
    rhs = (0, 1, 2, 3, 4, 5, 7) if not quick else (1, 3, 4)
    for p in "ds":
        for lay, uplo in itertools.product("CR", "LU"):
            for n in sizes + ((24, 33) if not quick else ()):
                yield "potrf", p, {"layout": lay, "uplo": uplo}, {"n": n}
                yield "sytrfnp", p, {"layout": lay, "uplo": uplo}, {"n": n}
                for nrhs in rhs:
                    if n > 17 and nrhs not in (1, 4):
                        continue
                    for rt in ("posv", "potrs", "sysvnp", "sytrsnp"):
                        yield rt, p, {"layout": lay, "uplo": uplo}, {"n": n, "nrhs": nrhs}
        for lay, side, uplo, tr, diag in itertools.product("CR", "LR", "LU", "NTC", "UN"):
            for m, n in ((0, 3), (3, 0), (1, 1), (5, 3), (3, 5), (9, 4), (4, 9), (8, 7)) if not quick else ((5, 3), (9, 4)):
                yield "trsm", p, {"layout": lay, "side": side, "uplo": uplo, "trans": tr, "diag": diag}, {"m": m, "n": n, "alpha": 1.0}
            yield "trsm", p, {"layout": lay, "side": side, "uplo": uplo, "trans": tr, "diag": diag}, {"m": 5, "n": 4, "alpha": -0.75}
        for lay in "CR":
            for m, n in ((0, 2), (2, 0), (1, 1), (5, 3), (3, 5), (16, 8), (9, 9), (12, 17)) if not quick else ((5, 3), (3, 5), (16, 8)):
                yield "geqrf", p, {"layout": lay}, {"m": m, "n": n}
                for tr in "NTC":
                    for nrhs in (0, 1, 2, 3, 5) if not quick else (2,):
                        yield "gels", p, {"layout": lay, "trans": tr}, {"m": m, "n": n, "nrhs": nrhs}
            for m, n, k in ((0, 0, 0), (3, 3, 0), (5, 3, 3), (5, 5, 2), (9, 7, 7), (16, 8, 8), (12, 10, 6)):
                yield "orgqr", p, {"layout": lay}, {"m": m, "n": n, "k": k}
            for side, tr in itertools.product("LR", "NT"):
                for m, n, k in ((0, 3, 0), (5, 3, 2), (3, 5, 3), (9, 4, 4), (16, 2, 8), (4, 9, 4)):
                    nq = m if side == "L" else n
                    if k > nq:
                        continue
                    yield "ormqr", p, {"layout": lay, "side": side, "trans": tr}, {"m": m, "n": n, "k": k}
