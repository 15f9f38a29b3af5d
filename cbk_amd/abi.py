"""Argument tables of the cbk.h C ABI, shared by every ctypes binding.

Each routine maps to its argument kinds in order: 'c' a char option, 'i' an int,
'T' the scalar type (double for d*, float for s*), 'p' a buffer pointer. The
lists are the signatures of the reference's include/cbk.h:77-351 (the same
names and argument order) plus the library's own extensions (include/cbk_ext.h).
"""

import ctypes

ROUTINES = {
    # include/cbk.h:77-351 (reference)
    "geqrf": "ciipipii",
    "ormqr": "ccciiipippiii",
    "orgqr": "ciiipipii",
    "potrf": "ccipiii",
    "potrs": "cciipipiii",
    "posv": "cciipipiii",
    "sytrfnp": "ccipiii",
    "sytrsnp": "cciipipiii",
    "sysvnp": "cciipipiii",
    "trsm": "cccccii" + "T" + "pipiii",
    "gels": "cciiipipipii",
}

# entry points of include/cbk_ext.h beyond the _info variants ('z' = size_t)
EXT = {
    "dsposv_compact": "cciipipipipzppii",
    "dsgels_compact": "cciiipipipippii",
    "dsposv_compact_noalloc": "cciipipipippii",
    "dssysvnp_compact": "cciipipipipzppii",
}
# dense <-> compact ('T' is the routine's scalar: cbk_s* float, else double)
for _p in "ds":
    EXT[f"cbk_{_p}pack_compact"] = "ciipcizpiTii"
    EXT[f"cbk_{_p}unpack_compact"] = "ciipipcizii"
    EXT[f"cbk_{_p}gepack_compact"] = "ciipipiii"
    EXT[f"cbk_{_p}geunpack_compact"] = "ciipipiii"

# the MKL-style adapter of include/cbk_compat.h (void; 'h' = the host int* info)
COMPAT = {
    "geqrf": "iiipippihii",
    "ormqr": "icciiipippipihii",
    "orgqr": "iiiipippihii",
    "potrf": "iiipihii",
    "sytrfnp": "iiipihii",
    "potrs": "iiiipipihii",
    "posv": "iiiipipihii",
    "sytrsnp": "iiiipipihii",
    "sysvnp": "iiiipipihii",
    "trsm": "iiiiiii" + "T" + "pipiii",
    "gels": "iciiipipipihii",
}
# MKL's enum values (include/cbk_compat.h)
MKL = {"ROW_MAJOR": 101, "COL_MAJOR": 102, "NOTRANS": 111, "TRANS": 112, "CONJTRANS": 113,
       "UPPER": 121, "LOWER": 122, "NONUNIT": 131, "UNIT": 132, "LEFT": 141, "RIGHT": 142,
       "COMPACT_SSE": 181, "COMPACT_AVX": 182, "COMPACT_AVX512": 183, "COMPACT_V64": 190}

# per-matrix info variants: the plain argument list plus a trailing device int[nm]
INFO_ROUTINES = ("potrf", "posv", "sytrfnp", "sysvnp", "gels")


def ctype_of(kind, scalar):
    if kind == "c":
        return ctypes.c_char
    if kind == "i":
        return ctypes.c_int
    if kind == "p":
        return ctypes.c_void_p
    if kind == "z":
        return ctypes.c_size_t
    if kind == "h":
        return ctypes.POINTER(ctypes.c_int)
    if kind == "T":
        return ctypes.c_double if scalar == "d" else ctypes.c_float
    raise ValueError(kind)


def bind(lib, name, kinds, scalar, restype=ctypes.c_int):
    fn = getattr(lib, name)
    fn.argtypes = [ctype_of(k, scalar) for k in kinds]
    fn.restype = restype
    return fn


def to_c(kind, value):
    if kind == "c":
        return value.encode() if isinstance(value, str) else value
    return value
