"""Test helpers: compact packing, seeded generators, and ctypes bindings of the
oracle (oracle/liboracle.so) and of the reference built from its sources
(oracle/_ref/libcbkref_v*.so, present only where the reference was compiled).

Compact layout (reference include/cbk.h:22-29): element (i,j) of matrix
idx = g*V + v sits at buf[g*ld*ncol*V + (j*ld + i)*V + v] column-major, with
i and j swapped row-major. Padded slots (idx >= nm) carry pad_diag on the
diagonal (tests/test_compact_util.hpp:440-475 of the reference).
"""

import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from cbk_amd import abi  # noqa: E402

ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

# value the tests put in the leading-dimension padding rows: a routine must
# neither read it into a result nor overwrite it
SENTINEL = 7777.0

DT = {"d": np.float64, "s": np.float32}


# ---------------------------------------------------------------- packing


def buf_shape(nm, rows, cols, ld, V, row=False):
    ng = (nm + V - 1) // V
    return (ng, rows, ld, V) if row else (ng, cols, ld, V)


def pack(M, V, ld=None, row=False, pad_diag=1.0, dtype=None, fill=SENTINEL):
    """Dense batch M[nm, rows, cols] -> flat compact buffer."""
    M = np.asarray(M)
    dtype = dtype or M.dtype
    nm, r, c = M.shape
    if ld is None:
        ld = c if row else r
    ng = (nm + V - 1) // V
    full = np.zeros((ng * V, r, c), dtype=dtype)
    full[:nm] = M
    for d in range(min(r, c)):
        full[nm:, d, d] = pad_diag
    F = full.reshape(ng, V, r, c)
    buf = np.full(buf_shape(nm, r, c, ld, V, row), fill, dtype=dtype)
    if row:
        buf[:, :, :c, :] = F.transpose(0, 2, 3, 1)
    else:
        buf[:, :, :r, :] = F.transpose(0, 3, 2, 1)
    return buf.reshape(-1)


def unpack(buf, nm, rows, cols, V, ld=None, row=False, all_slots=False):
    """Flat compact buffer -> dense batch [nm (or ng*V), rows, cols]."""
    if ld is None:
        ld = cols if row else rows
    ng = (nm + V - 1) // V
    b = np.asarray(buf).reshape(buf_shape(nm, rows, cols, ld, V, row))
    if row:
        F = b[:, :, :cols, :].transpose(0, 3, 1, 2)
    else:
        F = b[:, :, :rows, :].transpose(0, 3, 2, 1)
    F = F.reshape(ng * V, rows, cols)
    return F if all_slots else F[:nm]


def pack_vec(t, V, dtype=None):
    """Per-matrix vectors t[nm, k] (tau) -> compact (ld = k), padded with 0."""
    t = np.asarray(t)
    nm, k = t.shape
    ng = (nm + V - 1) // V
    full = np.zeros((ng * V, k), dtype=dtype or t.dtype)
    full[:nm] = t
    return full.reshape(ng, V, k).transpose(0, 2, 1).reshape(-1).copy()


def unpack_vec(buf, nm, k, V):
    ng = (nm + V - 1) // V
    return np.asarray(buf).reshape(ng, k, V).transpose(0, 2, 1).reshape(ng * V, k)[:nm]


# ---------------------------------------------------------------- generators


def rng(seed):
    return np.random.default_rng(seed)


def gen_uniform(r, nm, rows, cols, dtype=np.float64):
    return r.uniform(-1.0, 1.0, size=(nm, rows, cols)).astype(dtype)


def gen_spd(r, nm, n, dtype=np.float64, shift=None):
    """A = G G^T + shift*I (shift defaults to n), symmetric to the bit."""
    G = r.standard_normal((nm, n, n))
    A = G @ G.transpose(0, 2, 1) + (n if shift is None else shift) * np.eye(n)
    A = np.tril(A) + np.tril(A, -1).transpose(0, 2, 1)
    return A.astype(dtype)


def gen_sym_ldlt(r, nm, n, dtype=np.float64):
    """Symmetric indefinite with a known-good unpivoted LDL^T (as the reference's
    gen_sym_ldlt, tests/test_compact_util.hpp:376-399)."""
    L = np.tril(0.5 * r.uniform(-1, 1, (nm, n, n)), -1) + np.eye(n)
    mag = 0.5 + 2.0 * np.abs(r.uniform(-1, 1, (nm, n)))
    sgn = np.where(r.uniform(-1, 1, (nm, n)) < 0, -1.0, 1.0)
    if n >= 2:
        sgn[:, 1] = -1.0
    D = mag * sgn
    A = np.einsum("bik,bk,bjk->bij", L, D, L)
    A = np.tril(A) + np.tril(A, -1).transpose(0, 2, 1)
    return A.astype(dtype)


def gen_sym_dominant(r, nm, n, dtype=np.float64):
    """Symmetric, off-diagonals N(0,1), diagonal +/-(n+2) with random sign."""
    A = r.standard_normal((nm, n, n))
    A = np.tril(A, -1)
    A = A + A.transpose(0, 2, 1)
    sgn = np.where(r.uniform(-1, 1, (nm, n)) < 0, -1.0, 1.0)
    idx = np.arange(n)
    A[:, idx, idx] = sgn * (n + 2)
    return A.astype(dtype)


def gen_tri(r, nm, s, upper, dtype=np.float64):
    """Triangular with a diagonal bounded away from 0 (the reference's gen_tri,
    tests/test_compact_util.hpp:423-434); the other triangle zero."""
    A = r.uniform(-1, 1, (nm, s, s))
    A = np.triu(A) if upper else np.tril(A)
    d = np.diagonal(A, axis1=1, axis2=2)
    newd = np.where(d >= 0, 1.0, -1.0) * (2.0 + np.abs(r.uniform(-1, 1, (nm, s))))
    idx = np.arange(s)
    A[:, idx, idx] = newd
    return A.astype(dtype)


# ---------------------------------------------------------------- ctypes


class CLib:
    """ctypes binding of a library exporting the cbk.h routines, optionally
    with a name prefix (the oracle's oracle_)."""

    def __init__(self, path, prefix=""):
        self.path = path
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix
        self._fns = {}

    def fn(self, scalar, routine, info=False):
        key = (scalar, routine, info)
        if key not in self._fns:
            kinds = abi.ROUTINES[routine] + ("p" if info else "")
            name = f"{self.prefix}{scalar}{routine}_compact" + ("_info" if info else "")
            self._fns[key] = (abi.bind(self.lib, name, kinds, scalar), kinds)
        return self._fns[key]

    def call_raw(self, name, *args):
        """an include/cbk_ext.h entry point listed in abi.EXT, numpy buffers"""
        key = ("ext", name)
        if key not in self._fns:
            kinds = abi.EXT[name]
            scalar = "s" if name.startswith("cbk_s") else "d"
            self._fns[key] = (abi.bind(self.lib, self.prefix + name, kinds, scalar), kinds)
        f, kinds = self._fns[key]
        assert len(args) == len(kinds), (name, len(args), len(kinds))
        cargs = []
        for k, a in zip(kinds, args):
            if k == "p":
                cargs.append(None if a is None else a.ctypes.data_as(ctypes.c_void_p))
            else:
                cargs.append(abi.to_c(k, a))
        return f(*cargs)

    def call(self, scalar, routine, *args, info=None):
        """Call with numpy arrays for the buffers (passed by pointer, in place)."""
        f, kinds = self.fn(scalar, routine, info is not None)
        cargs = []
        full = list(args) + ([info] if info is not None else [])
        assert len(full) == len(kinds), (routine, len(full), len(kinds))
        for k, a in zip(kinds, full):
            if k == "p":
                cargs.append(None if a is None else a.ctypes.data_as(ctypes.c_void_p))
            else:
                cargs.append(abi.to_c(k, a))
        return f(*cargs)


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built: run __graft_entry__.build()")
        _oracle = CLib(ORACLE_SO, prefix="oracle_")
    return _oracle


def host_has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
        return " avx512f " in flags and " avx512vl " in flags and " avx512bw " in flags
    except OSError:
        return False


def ref_path():
    """The reference built here for this host's ISA, or None."""
    v = "v4" if host_has_avx512() else "v3"
    p = os.path.join(REF_DIR, f"libcbkref_{v}.so")
    return p if os.path.exists(p) else None


def ref():
    global _ref
    if _ref is None:
        p = ref_path()
        if p is None:
            return None
        _ref = CLib(p)
    return _ref
