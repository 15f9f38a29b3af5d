"""Dense <-> compact conversion on the GPU (cbk_?pack_compact,
cbk_?unpack_compact and the pointer-array cbk_?gepack_compact /
cbk_?geunpack_compact) against the oracle's restatement of the reference's
pack_compact / unpack_compact, bit for bit: every V, both layouts, a dense
side in the other layout, padded leading dimensions and batch strides,
partial groups, both pad values. The compact buffer's leading-dimension
padding and the dense buffer's padding must be left untouched (they hold a
sentinel). At full size (131072 FP64 32 x 32 matrices, V = 64) the pack is
checked against torch's own permutation of the batch and the unpack against
the input."""

import numpy as np
import pytest

import cbkt
import packcase

pytestmark = pytest.mark.gpu


def _gpu(name, *args):
    import cbk_amd
    import gpuutil
    dev = []
    targs = []
    for a in args:
        if isinstance(a, np.ndarray):
            d = gpuutil.to_dev(a)
            dev.append((a, d))
            targs.append(d)
        else:
            targs.append(a)
    st = cbk_amd.call_ext(name, *targs)
    assert cbk_amd.synchronize() == 0, cbk_amd.last_error()
    for a, d in dev:
        a[...] = d.cpu().numpy()
    return st


@pytest.mark.parametrize("case", packcase.CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_pack_unpack_match_oracle(case, dtype):
    layout, alayout, rows, cols, *_, V, nm = case
    print(f"case pack {case} {np.dtype(dtype).name}", flush=True)
    M, lda, ldap, stride = packcase.make(case, dtype, 21)
    p = "d" if dtype == np.float64 else "s"
    a = packcase.dense_batch(M, alayout, lda, stride)
    shape = cbkt.buf_shape(nm, rows, cols, ldap, V, layout == "R")
    for pad in (1.0, 0.0):
        want = np.full(shape, cbkt.SENTINEL, dtype=dtype).reshape(-1)
        got = want.copy()
        args = (layout, rows, cols, a, alayout, lda, stride)
        assert cbkt.oracle().call_raw(f"cbk_{p}pack_compact", *args, want, ldap, pad, V, nm) == 0
        assert _gpu(f"cbk_{p}pack_compact", *args, got, ldap, pad, V, nm) == 0
        assert np.array_equal(got, want)
    w = np.full_like(a, cbkt.SENTINEL)
    assert _gpu(f"cbk_{p}unpack_compact", layout, rows, cols, got, ldap, w, alayout, lda, stride,
                V, nm) == 0
    assert np.array_equal(w, a)


@pytest.mark.parametrize("case", [c for c in packcase.CASES if c[0] == c[1]])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_pointer_array_forms_match_oracle(case, dtype):
    import gpuutil
    layout, alayout, rows, cols, *_, V, nm = case
    print(f"case gepack {case} {np.dtype(dtype).name}", flush=True)
    M, lda, ldap, stride = packcase.make(case, dtype, 22)
    p = "d" if dtype == np.float64 else "s"
    torch = gpuutil.torch()
    a = packcase.dense_batch(M, alayout, lda, stride)
    da = gpuutil.to_dev(a)
    es = a.itemsize
    ptrs = torch.tensor([da.data_ptr() + es * k * stride for k in range(nm)] or [0],
                        dtype=torch.int64, device="cuda")
    shape = cbkt.buf_shape(nm, rows, cols, ldap, V, layout == "R")
    want = np.full(shape, cbkt.SENTINEL, dtype=dtype).reshape(-1)
    assert cbkt.oracle().call_raw(f"cbk_{p}pack_compact", layout, rows, cols, a, alayout, lda,
                                  stride, want, ldap, 1.0, V, nm) == 0
    dap = gpuutil.to_dev(np.full_like(want, cbkt.SENTINEL))
    import cbk_amd
    assert cbk_amd.call_ext(f"cbk_{p}gepack_compact", layout, rows, cols, ptrs, lda, dap, ldap, V,
                            nm) == 0
    assert cbk_amd.synchronize() == 0
    assert np.array_equal(dap.cpu().numpy(), want)
    dw = gpuutil.to_dev(np.full_like(a, cbkt.SENTINEL))
    wptrs = torch.tensor([dw.data_ptr() + es * k * stride for k in range(nm)] or [0],
                         dtype=torch.int64, device="cuda")
    assert cbk_amd.call_ext(f"cbk_{p}geunpack_compact", layout, rows, cols, wptrs, lda, dap, ldap,
                            V, nm) == 0
    assert cbk_amd.synchronize() == 0
    assert np.array_equal(dw.cpu().numpy(), a)


def test_pack_full_size_round_trip():
    """131072 FP64 32 x 32 matrices at V = 64: the compact buffer equals torch's
    permutation of the batch and unpacking restores it exactly"""
    import cbk_amd
    import gpuutil
    torch = gpuutil.torch()
    nm, n, V = 131072, 32, 64
    print("case pack full size", flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    M = torch.randn(nm, n, n, dtype=torch.float64, device="cuda", generator=g)
    # column-major dense matrices: element (i, j) of matrix k at k*n*n + j*n + i
    dense = M.transpose(1, 2).contiguous()
    ap = torch.empty(nm * n * n, dtype=torch.float64, device="cuda")
    assert cbk_amd.call_ext("cbk_dpack_compact", "C", n, n, dense, "C", n, n * n, ap, n, 1.0, V,
                            nm) == 0
    want = M.view(nm // V, V, n, n).permute(0, 3, 2, 1).reshape(-1)
    assert cbk_amd.synchronize() == 0
    assert torch.equal(ap, want)
    back = torch.full_like(dense, cbkt.SENTINEL)
    assert cbk_amd.call_ext("cbk_dunpack_compact", "C", n, n, ap, n, back, "C", n, n * n, V,
                            nm) == 0
    assert cbk_amd.synchronize() == 0
    assert torch.equal(back, dense)
