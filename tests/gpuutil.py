"""Run cbk.h calls on the GPU with numpy buffers: upload, call the HIP library
through its C ABI, synchronize, download in place. Mirrors CLib.call of the
oracle so one case runs through either."""

import numpy as np

import cbk_amd


_ready = False


def torch():
    global _ready
    import torch as _t
    if not _ready:
        _t.cuda.init()
        _t.zeros(1, device="cuda")  # a live context before the first library call
        _ready = True
    return _t


def to_dev(a):
    t = torch()
    return t.from_numpy(np.ascontiguousarray(a)).to("cuda")


def gpu_call(p, routine, *args, info=None):
    dev = []
    targs = []
    for a in args:
        if isinstance(a, np.ndarray):
            d = to_dev(a)
            dev.append((a, d))
            targs.append(d)
        else:
            targs.append(a)
    dinfo = None
    if info is not None:
        dinfo = to_dev(np.full(info.shape, -99, dtype=np.int32))
    st = cbk_amd.call(p, routine, *targs, info=dinfo)
    assert cbk_amd.synchronize() == 0, cbk_amd.last_error()
    for a, d in dev:
        a[...] = d.cpu().numpy()
    if info is not None:
        info[...] = dinfo.cpu().numpy()
    return st
