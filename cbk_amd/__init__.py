"""cbk_amd: compact batch kernels on MI355X.

The product is libcbk_amd.so, whose C ABI is include/cbk.h (the drop-in for
the reference's portable API, over device pointers) and include/cbk_ext.h.
This module binds that ABI with ctypes so Python callers (the tests, bench.py)
can pass torch CUDA tensors. It raises when the library is missing: there is
no CPU fallback.
"""

import ctypes
import os

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcbk_amd.so")

_lib = None


class LibraryMissing(RuntimeError):
    pass


def lib():
    """The loaded libcbk_amd.so (ctypes.CDLL)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                f"{LIB_PATH} is not built; run `python -m cbk_amd.build` "
                "(or __graft_entry__.build())")
        # One HIP runtime per process: PyTorch ships its own libamdhip64.so.7,
        # and loading this library first would pull /opt/rocm's copy in beside
        # it (two runtimes, two device views). With torch loaded first the
        # dynamic linker resolves this library's libamdhip64.so.7 to torch's.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.cbk_set_stream.argtypes = [ctypes.c_void_p]
        _lib.cbk_set_stream.restype = ctypes.c_int
        _lib.cbk_get_stream.argtypes = []
        _lib.cbk_get_stream.restype = ctypes.c_void_p
        _lib.cbk_synchronize.argtypes = []
        _lib.cbk_synchronize.restype = ctypes.c_int
        _lib.cbk_set_policy.argtypes = [ctypes.c_int]
        _lib.cbk_set_policy.restype = ctypes.c_int
        _lib.cbk_last_error_string.argtypes = []
        _lib.cbk_last_error_string.restype = ctypes.c_char_p
        _lib.cbk_get_version.argtypes = []
        _lib.cbk_get_version.restype = ctypes.c_int
    return _lib


_fns = {}


def fn(scalar, routine, variant=""):
    """ctypes function for {scalar}{routine}_compact{variant} with argtypes set.
    variant: "" (cbk.h) or "_info" (cbk_ext.h, trailing int* info)."""
    key = (scalar, routine, variant)
    if key not in _fns:
        kinds = abi.ROUTINES[routine] + ("p" if variant == "_info" else "")
        name = f"{scalar}{routine}_compact{variant}"
        _fns[key] = (abi.bind(lib(), name, kinds, scalar), kinds)
    return _fns[key]


def _ptr(x):
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        return ctypes.c_void_p(x.data_ptr())
    if isinstance(x, int):
        return ctypes.c_void_p(x)
    raise TypeError(f"expected a device tensor or address, got {type(x)}")


def call(scalar, routine, *args, info=None, variant=None):
    """Call a routine with torch CUDA tensors (or raw device addresses) for the
    buffers; returns the routine's int status. Launches on the current cbk
    stream (see set_stream)."""
    if variant is None:
        variant = "_info" if info is not None else ""
    f, kinds = fn(scalar, routine, variant)
    full = list(args) + ([info] if variant == "_info" else [])
    if len(full) != len(kinds):
        raise TypeError(f"{scalar}{routine}: expected {len(kinds)} arguments, got {len(full)}")
    cargs = [_ptr(a) if k == "p" else abi.to_c(k, a) for k, a in zip(kinds, full)]
    return f(*cargs)


def call_ext(name, *args):
    """Call an include/cbk_ext.h entry point listed in abi.EXT (device tensors
    or addresses for its pointers)."""
    key = ("ext", name, "")
    if key not in _fns:
        kinds = abi.EXT[name]
        _fns[key] = (abi.bind(lib(), name, kinds, "s" if name.startswith("cbk_s") else "d"), kinds)
    f, kinds = _fns[key]
    if len(args) != len(kinds):
        raise TypeError(f"{name}: expected {len(kinds)} arguments, got {len(args)}")
    return f(*[_ptr(a) if k == "p" else abi.to_c(k, a) for k, a in zip(kinds, args)])


def call_compat(scalar, routine, *args):
    """Call cbk_{scalar}{routine}_compact of include/cbk_compat.h (the
    MKL-style adapter) with device tensors or addresses for its buffers and None
    in the info slot; returns the host info it wrote (None for trsm, which has
    none)."""
    key = ("compat", scalar, routine)
    if key not in _fns:
        kinds = abi.COMPAT[routine]
        _fns[key] = (abi.bind(lib(), f"cbk_{scalar}{routine}_compact", kinds, scalar, None), kinds)
    f, kinds = _fns[key]
    if len(args) != len(kinds):
        raise TypeError(f"cbk_{scalar}{routine}_compact: expected {len(kinds)} arguments, got {len(args)}")
    info = ctypes.c_int(-99)
    cargs = []
    for k, a in zip(kinds, args):
        if k == "h":
            cargs.append(ctypes.pointer(info))
        elif k == "p":
            cargs.append(_ptr(a))
        else:
            cargs.append(abi.to_c(k, a))
    f(*cargs)
    return info.value if "h" in kinds else None


def dsposv_worksize(layout, uplo, n, nrhs, ldap, ldbp, ldxp, V, nm):
    f = lib().dsposv_compact_worksize
    f.argtypes = [ctypes.c_char, ctypes.c_char] + [ctypes.c_int] * 7
    f.restype = ctypes.c_size_t
    return f(abi.to_c("c", layout), abi.to_c("c", uplo), n, nrhs, ldap, ldbp, ldxp, V, nm)


def set_stream(stream):
    """Make `stream` (a torch.cuda.Stream, a raw hipStream_t int, or None) the
    stream this thread's calls launch on."""
    if stream is None:
        h = None
    elif isinstance(stream, int):
        h = stream
    else:
        h = stream.cuda_stream
    return lib().cbk_set_stream(h)


POLICY_AUTO = 0
POLICY_LANE = 1


def set_policy(policy):
    """POLICY_AUTO (tuned kernels where they apply) or POLICY_LANE (the general
    lane kernels everywhere); returns the previous policy."""
    return lib().cbk_set_policy(policy)


def synchronize():
    return lib().cbk_synchronize()


def last_error():
    """the HIP runtime's message for this thread's last CBK_ERR_LAUNCH"""
    return lib().cbk_last_error_string().decode()


def version():
    return lib().cbk_get_version()


# ---- uneven batches (include/cbk_ext.h: cbk_gels_batch_plan, dgels_batch, ...)

class GelsProblem(ctypes.Structure):
    """cbk_gels_problem: sizes, leading dimensions and element offsets of one
    column-major problem of an uneven batch"""
    _fields_ = [("m", ctypes.c_int), ("n", ctypes.c_int), ("nrhs", ctypes.c_int),
                ("lda", ctypes.c_int), ("ldb", ctypes.c_int),
                ("a_off", ctypes.c_longlong), ("b_off", ctypes.c_longlong)]


BATCH_NCLASS = 7


def problem_dtype():
    """numpy dtype laid out as cbk_gels_problem, for building large batches
    without a Python loop"""
    import numpy as np
    return np.dtype([("m", "<i4"), ("n", "<i4"), ("nrhs", "<i4"), ("lda", "<i4"),
                     ("ldb", "<i4"), ("a_off", "<i8"), ("b_off", "<i8")], align=True)


class Tuning(ctypes.Structure):
    """cbk_tuning"""
    _fields_ = [("narrow_max_cols", ctypes.c_int * 3),
                ("blocks_per_cu", ctypes.c_int * BATCH_NCLASS)]


def _bind_batch():
    L = lib()
    if getattr(L, "_batch_bound", False):
        return L
    P = ctypes.c_void_p
    L.cbk_gels_batch_plan.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int,
                                      ctypes.POINTER(GelsProblem), ctypes.POINTER(P),
                                      ctypes.POINTER(ctypes.c_int)]
    L.cbk_gels_batch_plan.restype = ctypes.c_int
    L.cbk_batch_plan_worksize.argtypes = [P]
    L.cbk_batch_plan_worksize.restype = ctypes.c_size_t
    L.cbk_batch_plan_classes.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
    L.cbk_batch_plan_classes.restype = ctypes.c_int
    for name in ("dgels_batch", "sgels_batch"):
        f = getattr(L, name)
        f.argtypes = [P, P, P, P, P, ctypes.c_size_t]
        f.restype = ctypes.c_int
    L.dsgels_batch.argtypes = [P, P, P, P, P, P, P, ctypes.c_size_t]
    L.dsgels_batch.restype = ctypes.c_int
    L.cbk_batch_plan_destroy.argtypes = [P]
    L.cbk_batch_plan_destroy.restype = ctypes.c_int
    L.cbk_autotune.argtypes = []
    L.cbk_autotune.restype = ctypes.c_int
    L.cbk_get_tuning.argtypes = [ctypes.POINTER(Tuning)]
    L.cbk_get_tuning.restype = ctypes.c_int
    L.cbk_set_tuning.argtypes = [ctypes.POINTER(Tuning)]
    L.cbk_set_tuning.restype = ctypes.c_int
    L._batch_bound = True
    return L


class GelsBatchPlan:
    """A plan over an uneven batch (cbk_gels_batch_plan). problems: a sequence
    of (m, n, nrhs, lda, ldb, a_off, b_off). Raises ValueError with the code
    (and the index of a bad descriptor) on invalid input."""

    def __init__(self, prec, trans, problems):
        L = _bind_batch()
        if hasattr(problems, "dtype"):  # numpy array of PROBLEM_DTYPE
            import numpy as np
            assert problems.dtype == problem_dtype() and problems.flags.c_contiguous
            arr = problems.ctypes.data_as(ctypes.POINTER(GelsProblem))
            self._keep = problems
        else:
            arr = (GelsProblem * max(1, len(problems)))()
            for i, p in enumerate(problems):
                arr[i] = GelsProblem(*[int(x) for x in p])
        self._h = ctypes.c_void_p()
        bad = ctypes.c_int(-1)
        st = L.cbk_gels_batch_plan(prec.encode(), trans.encode(), len(problems), arr,
                                   ctypes.byref(self._h), ctypes.byref(bad))
        if st != 0:
            raise ValueError(f"cbk_gels_batch_plan returned {st} (bad descriptor {bad.value})")
        self.prec = prec
        self.worksize = L.cbk_batch_plan_worksize(self._h)

    def classes(self):
        cnt = (ctypes.c_int * BATCH_NCLASS)()
        _bind_batch().cbk_batch_plan_classes(self._h, cnt)
        return list(cnt)

    def run(self, ap, bp, info, work, lwork=None):
        """dgels_batch / sgels_batch on device tensors (info and work may be
        None); lwork defaults to work's size in bytes"""
        f = _bind_batch().dgels_batch if self.prec == "D" else _bind_batch().sgels_batch
        if lwork is None:
            lwork = work.numel() * work.element_size() if hasattr(work, "numel") else 0
        return f(self._h, _ptr(ap), _ptr(bp), _ptr(info), _ptr(work), lwork)

    def run_mixed(self, ap, bp, xp, iters, info, work, lwork=None):
        """dsgels_batch (a plan made with prec 'M') on device tensors"""
        if lwork is None:
            lwork = work.numel() * work.element_size() if hasattr(work, "numel") else 0
        return _bind_batch().dsgels_batch(self._h, _ptr(ap), _ptr(bp), _ptr(xp), _ptr(iters),
                                          _ptr(info), _ptr(work), lwork)

    def close(self):
        if self._h:
            _bind_batch().cbk_batch_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune():
    return _bind_batch().cbk_autotune()


FAMILY = {"default": 0, "register": 1, "quad": 2, "team": 3, "columns": 4, "wave": 5, "lane": 6,
          "grid": 7}


def _bind_tune():
    L = lib()
    if not getattr(L, "_tune_bound", False):
        key = [ctypes.c_char, ctypes.c_char_p, ctypes.c_char, ctypes.c_char] + [ctypes.c_int] * 4
        L.cbk_autotune_compact.argtypes = key + [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                 ctypes.POINTER(ctypes.c_int)]
        L.cbk_autotune_compact.restype = ctypes.c_int
        L.cbk_set_kernel_family.argtypes = key + [ctypes.c_int]
        L.cbk_set_kernel_family.restype = ctypes.c_int
        L.cbk_get_kernel_family.argtypes = key
        L.cbk_get_kernel_family.restype = ctypes.c_int
        L._tune_bound = True
    return L


def _key(prec, routine, uplo, trans, m, n, nrhs, V):
    return (abi.to_c("c", prec), routine.encode(), abi.to_c("c", uplo), abi.to_c("c", trans),
            m, n, nrhs, V)


def autotune_compact(prec, routine, m, n, nrhs, V, nm=0, uplo="L", trans="N"):
    """cbk_autotune_compact: (status, family installed, V chosen)"""
    f, v = ctypes.c_int(-1), ctypes.c_int(-1)
    st = _bind_tune().cbk_autotune_compact(*_key(prec, routine, uplo, trans, m, n, nrhs, V), nm,
                                           ctypes.byref(f), ctypes.byref(v))
    return st, f.value, v.value


def set_kernel_family(prec, routine, m, n, nrhs, V, family, uplo="L", trans="N"):
    return _bind_tune().cbk_set_kernel_family(*_key(prec, routine, uplo, trans, m, n, nrhs, V),
                                              family)


def get_kernel_family(prec, routine, m, n, nrhs, V, uplo="L", trans="N"):
    return _bind_tune().cbk_get_kernel_family(*_key(prec, routine, uplo, trans, m, n, nrhs, V))


def get_tuning():
    t = Tuning()
    assert _bind_batch().cbk_get_tuning(ctypes.byref(t)) == 0
    return {"narrow_max_cols": list(t.narrow_max_cols), "blocks_per_cu": list(t.blocks_per_cu)}


def set_tuning(narrow_max_cols, blocks_per_cu):
    t = Tuning()
    t.narrow_max_cols[:] = list(narrow_max_cols)
    t.blocks_per_cu[:] = list(blocks_per_cu)
    return _bind_batch().cbk_set_tuning(ctypes.byref(t))
