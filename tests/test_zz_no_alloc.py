"""dsposv_compact_noalloc (the entry point that allocates nothing) against
dsposv_compact with a caller-allocated workspace, through the C ABI.

Both run the same kernels on the same inputs, so every output buffer must be
identical bit for bit: X, iter, info, and A (untouched on success, the FP64
factor on the fallback). Two calls in a row, the second smaller, on the lane
path (whose workspace the static array provides), then a fused-kernel shape
(no workspace), the size limit (CBK_ERR_SCRATCH, nothing launched) and the
argument codes."""

import numpy as np
import pytest

import cbkt

pytestmark = pytest.mark.gpu

CBK_ERR_SCRATCH = -1001


def spd(r, nm, n):
    G = r.standard_normal((nm, n, n))
    A = G @ np.swapaxes(G, 1, 2) + n * np.eye(n)
    return 0.5 * (A + np.swapaxes(A, 1, 2))


def run(name, A, B, V, uplo="L"):
    import cbk_amd
    import gpuutil
    torch = gpuutil.torch()
    nm, n, _ = A.shape
    nrhs = B.shape[2]
    ap = cbkt.pack(np.where(np.eye(n, dtype=bool), A, np.tril(A) if uplo == "L" else np.triu(A)), V)
    bp = cbkt.pack(B, V, pad_diag=0.0)
    dA, dB = gpuutil.to_dev(ap), gpuutil.to_dev(bp)
    dX = gpuutil.to_dev(np.full_like(bp, -5.0))
    dI = gpuutil.to_dev(np.full(nm, -77, np.int32))
    dN = gpuutil.to_dev(np.full(nm, -77, np.int32))
    if name == "noalloc":
        st = cbk_amd.call_ext("dsposv_compact_noalloc", "C", uplo, n, nrhs, dA, n, dB, n, dX, n, dI,
                              dN, V, nm)
    else:
        ws = cbk_amd.dsposv_worksize("C", uplo, n, nrhs, n, n, n, V, nm)
        dW = torch.empty(max(ws, 8), dtype=torch.uint8, device="cuda") if ws else None
        st = cbk_amd.call_ext("dsposv_compact", "C", uplo, n, nrhs, dA, n, dB, n, dX, n, dW, ws, dI,
                              dN, V, nm)
    assert cbk_amd.synchronize() == 0, cbk_amd.last_error()
    return st, [d.cpu().numpy() for d in (dA, dX, dI, dN)]


def _same(a, b):
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("shape", [
    # (n, nrhs, V, nm) of the first call, then of the smaller second call
    ((72, 5, 8, 601), (72, 3, 8, 117)),    # the lane path (n > 64): workspace from the static array
    ((7, 2, 2, 999), (5, 1, 2, 33)),       # the lane path (V < 4)
    ((32, 4, 64, 300), (32, 2, 64, 70)),   # the fused kernel: no workspace at all
])
def test_noalloc_matches_workspace_call(shape):
    import cbk_amd
    assert cbk_amd.lib().cbk_noalloc_scratch_bytes() == 32 << 20
    r = cbkt.rng(515 + shape[0][0])
    for (n, nrhs, V, nm) in shape:
        A = spd(r, nm, n)
        A[3] = A[3] - 3.0 * np.eye(n) * np.abs(A[3]).max()  # not SPD: FP64 fallback fails too
        B = r.standard_normal((nm, n, nrhs))
        print(f"case noalloc n={n} nrhs={nrhs} V={V} nm={nm}", flush=True)
        s1, out1 = run("noalloc", A, B, V)
        s2, out2 = run("workspace", A, B, V)
        assert s1 == s2 == 0
        for a, b in zip(out1, out2):
            assert _same(a, b)
        info = out1[3]
        assert info[3] > 0 and np.count_nonzero(info) == 1


def test_noalloc_limit_and_codes():
    import cbk_amd
    z, fake = None, 8  # fake: a non-null address that is never dereferenced
    f = lambda *a: cbk_amd.call_ext("dsposv_compact_noalloc", *a)
    # a call the lane kernel serves (V < 4; the fused kernels need none and
    # would launch) needs groups * V * (n*n + n*nrhs) * 4 bytes of workspace,
    # here > 32 MiB: refused before any launch
    assert cbk_amd.dsposv_worksize("C", "L", 40, 5, 40, 40, 40, 2, 20000) > (32 << 20)
    assert f("C", "L", 40, 5, z, 40, z, 40, z, 40, fake, fake, 2, 20000) == CBK_ERR_SCRATCH
    assert f("X", "L", 40, 5, z, 40, z, 40, z, 40, fake, fake, 2, 20000) == -1
    assert f("C", "L", 40, 5, z, 40, z, 40, z, 40, z, fake, 2, 10) == -11
    assert f("C", "L", 40, 5, z, 40, z, 40, z, 40, fake, z, 2, 10) == -12
    assert f("C", "L", 40, 5, z, 40, z, 40, z, 40, fake, fake, 3, 10) == -13
    assert f("C", "L", 40, 5, z, 40, z, 40, z, 40, fake, fake, 2, -1) == -14
    assert f("C", "L", 40, 5, z, 40, z, 40, z, 40, z, z, 2, 0) == 0
