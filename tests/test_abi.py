"""The C-ABI library loads and exports exactly what include/*.h declares
(no GPU needed: nothing is launched). Also: argument validation, which returns
before any launch, against the oracle's (the reference's) -j codes."""

import ctypes
import itertools
import os
import re

import numpy as np
import pytest

import cbk_amd
import cbkt

INCLUDE = os.path.join(cbkt.ROOT, "include")


def declared_functions():
    names = set()
    for h in sorted(os.listdir(INCLUDE)):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*([a-z_0-9]+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def exported_functions():
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", cbk_amd.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def test_library_exports_every_declared_symbol():
    decl = declared_functions()
    assert "dposv_compact" in decl and "cbk_get_version" in decl and "dgels_compact_info" in decl
    lib = cbk_amd.lib()
    missing = [n for n in sorted(decl) if not hasattr(lib, n)]
    assert not missing, missing
    extra = exported_functions() - decl
    assert not extra, sorted(extra)


def test_reference_entry_points_all_present():
    """The 22 entry points of the reference's cbk.h plus cbk_get_version."""
    lib = cbk_amd.lib()
    for p in "ds":
        for r in cbk_amd.abi.ROUTINES:
            assert hasattr(lib, f"{p}{r}_compact")
    assert cbk_amd.version() == 200


def test_argument_codes_match_oracle():
    """Illegal arguments return the reference's -j before any launch; the
    pointers are never touched, so null pointers are safe here."""
    grids = {
        "geqrf": [("X", "C", "r"), (-1, 3), (-1, 0, 3), None, (0, 2, 3), None, (3, 16, 64, 128), (-1, 0)],
        "ormqr": [("X", "C", "R"), ("X", "L", "R"), ("X", "N", "T", "C"), (-1, 2), (-1, 2), (-1, 0, 3),
                  None, (0, 1, 2), None, None, (0, 2), (3, 8), (-1, 0)],
        "orgqr": [("X", "C"), (-1, 3), (-1, 2, 4), (-1, 1, 3), None, (0, 3), None, (5, 8), (-1, 0)],
        "potrf": [("X", "C", "R"), ("X", "L", "u"), (-1, 0, 3), None, (0, 2, 3), (1, 16, 64, 65), (-1, 0)],
        "sytrfnp": [("X", "c"), ("X", "l"), (-1, 3), None, (2, 3), (0, 8), (-1, 0)],
        "posv": [("X", "C", "R"), ("X", "L", "U"), (-1, 0, 3), (-1, 0, 2), None, (1, 3), None, (1, 2, 3),
                 (7, 8, 32), (-1, 0)],
        "potrs": [("X", "C"), ("Q", "L"), (-1, 3), (-1, 2), None, (2, 3), None, (1, 3), (8,), (-1, 0)],
        "sysvnp": [("X", "R"), ("X", "U"), (-1, 3), (-1, 2), None, (3,), None, (1, 2), (8, 9), (-1, 0)],
        "sytrsnp": [("C",), ("L", "Z"), (3,), (2,), None, (2, 3), None, (2, 3), (8,), (-1, 0)],
        "trsm": [("X", "C"), ("X", "L", "R"), ("X", "U"), ("X", "N", "C"), ("X", "U", "N"), (-1, 3), (-1, 2),
                 (2.0,), None, (1, 3), None, (1, 3), (8, 6), (-1, 0)],
        "gels": [("X", "C", "R"), ("X", "N", "T", "c"), (-1, 3), (-1, 2), (-1, 0, 2), None, (1, 3), None,
                 (1, 2, 3), None, (8, 7), (-1, 0)],
    }
    lib = cbk_amd.lib()
    o = cbkt.oracle()
    checked = 0
    for rt, grid in grids.items():
        kinds = cbk_amd.abi.ROUTINES[rt]
        choices = [g if g is not None else (None,) for g in grid]
        for combo in itertools.product(*choices):
            # only combinations the oracle rejects (nothing is launched then)
            big = np.zeros(1 << 12)
            args_o = [big.copy() if k == "p" else v for k, v in zip(kinds, combo)]
            so = o.call("d", rt, *args_o)
            if so == 0:
                continue
            sg = cbk_amd.call("d", rt, *[None if k == "p" else v for k, v in zip(kinds, combo)])
            assert sg == so, (rt, combo, sg, so)
            checked += 1
    assert checked > 500


def test_quick_returns_launch_nothing():
    """Empty problems return 0 without touching (null) device pointers."""
    assert cbk_amd.call("d", "posv", "C", "L", 0, 3, None, 1, None, 1, 64, 10) == 0
    assert cbk_amd.call("d", "posv", "C", "L", 4, 3, None, 4, None, 4, 64, 0) == 0
    assert cbk_amd.call("d", "potrs", "C", "L", 4, 0, None, 4, None, 4, 64, 10) == 0
    assert cbk_amd.call("d", "gels", "C", "N", 4, 3, 0, None, 4, None, 4, None, 64, 10) == 0
    assert cbk_amd.call("s", "geqrf", "C", 0, 3, None, 1, None, 64, 10) == 0
    assert cbk_amd.call("s", "ormqr", "C", "L", "T", 4, 3, 0, None, 4, None, None, 4, 64, 10) == 0
    assert cbk_amd.call("d", "trsm", "C", "L", "L", "N", "U", 0, 3, 1.0, None, 1, None, 1, 64, 5) == 0


def test_mixed_entry_points_reject_null_outputs():
    """dsposv_compact needs iter and info, dsgels_compact iter, when nm > 0:
    a NULL is an argument error (returned before any launch), not a fault"""
    z, fake = None, 8  # fake: a non-null address that is never dereferenced
    ds = lambda *a: cbk_amd.call_ext("dsposv_compact", *a)
    # (layout, uplo, n, nrhs, ap, ldap, bp, ldbp, xp, ldxp, work, lwork, iter, info, V, nm)
    assert ds("C", "L", 32, 4, z, 32, z, 32, z, 32, z, 0, z, fake, 64, 10) == -13
    assert ds("C", "L", 32, 4, z, 32, z, 32, z, 32, z, 0, fake, z, 64, 10) == -14
    assert ds("C", "L", 32, 4, z, 32, z, 32, z, 32, z, 0, z, z, 3, 10) == -13
    assert ds("C", "L", 32, 4, z, 32, z, 32, z, 32, z, 0, fake, fake, 3, 10) == -15
    assert ds("C", "L", 32, 4, z, 32, z, 32, z, 32, z, 0, z, z, 64, 0) == 0  # nothing to do
    # an ldap too large for the fused kernel's 32-bit offsets needs the lane path's workspace
    ws = cbk_amd.dsposv_worksize("C", "L", 32, 4, 1 << 20, 32, 32, 64, 10)
    assert ws == 64 * (32 * 32 + 32 * 4) * 4
    assert cbk_amd.dsposv_worksize("C", "L", 32, 4, 32, 32, 32, 64, 10) == 0
    assert ds("C", "L", 32, 4, z, 1 << 20, z, 32, z, 32, z, 0, fake, fake, 64, 10) == -12
    dg = lambda *a: cbk_amd.call_ext("dsgels_compact", *a)
    # (layout, trans, m, n, nrhs, ap, ldap, bp, ldbp, xp, ldxp, iter, info, V, nm)
    assert dg("C", "N", 64, 32, 2, z, 64, z, 64, z, 32, z, z, 64, 10) == -12
    assert dg("C", "N", 64, 32, 2, z, 64, z, 64, z, 32, z, z, 64, 0) == 0
