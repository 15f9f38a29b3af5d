"""Pin the oracle to the reference, bit for bit.

1. Against tests/golden/ref_vectors.npz: vectors the reference library (its own
   src/cbk.cpp, compiled by oracle/Makefile `ref`) wrote on small seeded cases
   of all 11 routines x both precisions x every option set. Runs anywhere.
2. Against the live reference build in oracle/_ref/ on a wider sweep
   (V = 2 and 16, padded last group), where that build exists.
"""

import json
import os

import numpy as np
import pytest

import cases
import cbkt

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")


def _load_golden():
    z = np.load(GOLDEN, allow_pickle=False)
    man = json.loads(bytes(z["manifest"]).decode())
    return z, man


def test_golden_fixture_present():
    z, man = _load_golden()
    routines = {c["routine"] for c in man["cases"]}
    assert routines == {"potrf", "sytrfnp", "posv", "potrs", "sysvnp", "sytrsnp", "trsm",
                        "geqrf", "gels", "orgqr", "ormqr"}
    assert {c["p"] for c in man["cases"]} == {"d", "s"}


def test_oracle_matches_reference_golden_vectors():
    z, man = _load_golden()
    o = cbkt.oracle()
    bad = []
    for c in man["cases"]:
        args = [z[a["array"]].copy() if "array" in a else a["value"] for a in c["args"]]
        st = o.call(c["p"], c["routine"], *args)
        outs = [args[i] for i in c["outs_idx"]]
        exp = [z[k] for k in c["outs"]]
        if st != c["status"] or not all(cases.same_bits(a, b) for a, b in zip(outs, exp)):
            bad.append((c["routine"], c["p"], c["opts"], c["dims"]))
    assert not bad, bad[:10]


def test_golden_vectors_are_discriminating():
    """A planted mistake in the oracle's output (one ulp on one element of each
    case that computed something) must be caught by the bitwise comparison."""
    z, man = _load_golden()
    o = cbkt.oracle()
    checked = 0
    for c in man["cases"][::7]:
        args = [z[a["array"]].copy() if "array" in a else a["value"] for a in c["args"]]
        o.call(c["p"], c["routine"], *args)
        out = args[c["outs_idx"][0]]
        if out.size == 0:
            continue
        exp = z[c["outs"][0]]
        out = out.copy()
        out.flat[0] = np.nextafter(out.flat[0], np.inf)
        assert not cases.same_bits(out, exp)
        checked += 1
    assert checked > 20


@pytest.mark.skipif(cbkt.ref() is None, reason="reference not built here (oracle/_ref)")
@pytest.mark.parametrize("V,nm", [(2, 5), (16, 35)])
def test_oracle_matches_live_reference(V, nm):
    o, r = cbkt.oracle(), cbkt.ref()
    bad = []
    n = 0
    for i, (rt, p, opts, dims) in enumerate(cases.sweep()):
        if i % 2:
            continue
        c = cases.make_case(rt, p, opts, dims, V, nm, seed=9000 + i, ld_extra=i % 2)
        so, outo = cases.run(o.call, c)
        sr, outr = cases.run(r.call, c)
        n += 1
        if so != sr or not all(cases.same_bits(a, b) for a, b in zip(outo, outr)):
            bad.append((rt, p, opts, dims))
    assert n > 1000
    assert not bad, bad[:10]


@pytest.mark.skipif(cbkt.ref() is None, reason="reference not built here (oracle/_ref)")
def test_argument_codes_match_reference():
    """Every -j code of every routine, against the reference's own validation
    (src/cbk.cpp:61-320), on a grid of illegal arguments."""
    o, r = cbkt.oracle(), cbkt.ref()
    dummy = np.zeros(1 << 16)
    dummys = np.zeros(1 << 16, dtype=np.float32)
    bad = []
    grids = {
        "geqrf": [("X", "C"), (-1, 0, 3), (-1, 0, 3), None, (0, 2, 3), None, (3, 4, 32, 64), (-1, 0, 2)],
        "ormqr": [("X", "C", "R"), ("X", "L", "R"), ("X", "N", "T", "C"), (-1, 2), (-1, 2), (-1, 0, 3),
                  None, (0, 1, 2), None, None, (0, 2), (3, 8), (-1, 2)],
        "orgqr": [("X", "C", "r"), (-1, 3), (-1, 2, 4), (-1, 1, 3), None, (0, 3), None, (5, 8), (-1, 2)],
        "potrf": [("X", "C", "R"), ("X", "L", "u"), (-1, 0, 3), None, (0, 2, 3), (1, 16, 64), (-1, 1)],
        "sytrfnp": [("X", "c"), ("X", "l"), (-1, 3), None, (2, 3), (0, 8), (-1, 1)],
        "posv": [("X", "C", "R"), ("X", "L", "U"), (-1, 0, 3), (-1, 0, 2), None, (1, 3), None, (1, 2, 3),
                 (7, 8), (-1, 0)],
        "potrs": [("X", "C"), ("Q", "L"), (-1, 3), (-1, 2), None, (2, 3), None, (1, 3), (8,), (-1, 0)],
        "sysvnp": [("X", "R"), ("X", "U"), (-1, 3), (-1, 2), None, (3,), None, (1, 2), (8, 9), (-1,)],
        "sytrsnp": [("C",), ("L", "Z"), (3,), (2,), None, (2, 3), None, (2, 3), (8,), (0,)],
        "trsm": [("X", "C"), ("X", "L", "R"), ("X", "U"), ("X", "N", "C"), ("X", "U", "N"), (-1, 3), (-1, 2),
                 (0.0, 2.0), None, (1, 3), None, (1, 3), (8, 6), (-1, 0)],
        "gels": [("X", "C", "R"), ("X", "N", "T", "c"), (-1, 3), (-1, 2), (-1, 0, 2), None, (1, 3), None,
                 (1, 2, 3), None, (8, 7), (-1, 1)],
    }
    import itertools
    for rt, grid in grids.items():
        kinds = cbkt.abi.ROUTINES[rt]
        choices = [g if g is not None else (None,) for g in grid]
        combos = list(itertools.product(*choices))
        for k, combo in enumerate(combos[:: max(1, len(combos) // 400)]):
            for p, buf in (("d", dummy), ("s", dummys)):
                args = [buf if kd == "p" else v for kd, v in zip(kinds, combo)]
                so = o.call(p, rt, *[a.copy() if isinstance(a, np.ndarray) else a for a in args])
                sr = r.call(p, rt, *[a.copy() if isinstance(a, np.ndarray) else a for a in args])
                # the oracle widens V to {32, 64}; everything else must agree
                vpos = {"geqrf": 6, "ormqr": 11, "orgqr": 7, "potrf": 5, "sytrfnp": 5, "posv": 8,
                        "potrs": 8, "sysvnp": 8, "sytrsnp": 8, "trsm": 12, "gels": 10}[rt]
                if combo[vpos] in (32, 64) and so != sr:
                    continue
                if so != sr:
                    bad.append((rt, p, combo, so, sr))
    assert not bad, bad[:10]
