"""GPU parity of every routine through the C ABI, against the oracle, bit for bit.

The oracle is pinned bit-for-bit to the reference (test_oracle_ref.py); the
HIP kernels perform the same operations in the same order, so their outputs
must equal the oracle's exactly: every output buffer (factors, tau, solutions,
the untouched opposite triangle, the leading-dimension padding and the padded
group slots), for both precisions, every option set, V in {2, 8, 64}, and
partial last groups.
"""

import numpy as np
import pytest

import cases
import cbkt

pytestmark = pytest.mark.gpu


def _gpu():
    import gpuutil
    return gpuutil


@pytest.fixture(autouse=True)
def lane_policy():
    """these tests pin the general lane kernels: force them for every shape"""
    import cbk_amd
    prev = cbk_amd.set_policy(cbk_amd.POLICY_LANE)
    yield
    cbk_amd.set_policy(prev)


def _check(case):
    so, outo = cases.run(cbkt.oracle().call, case)
    sg, outg = cases.run(_gpu().gpu_call, case)
    assert so == sg
    for a, b in zip(outo, outg):
        if not cases.same_bits(a, b):
            d = np.abs(a.astype(np.float64) - b.astype(np.float64))
            raise AssertionError(f"{case['routine']}{case['p']} {case['opts']} {case['dims']} "
                                 f"V={case['V']} nm={case['nm']}: max diff {np.nanmax(d)}")


ROUTINES = ["potrf", "sytrfnp", "posv", "potrs", "sysvnp", "sytrsnp", "trsm", "geqrf",
            "gels", "orgqr", "ormqr"]


@pytest.mark.parametrize("routine", ROUTINES)
@pytest.mark.parametrize("V,nm", [(2, 5), (8, 11), (32, 71), (64, 130)])
def test_lane_parity(routine, V, nm):
    n = 0
    for i, (rt, p, opts, dims) in enumerate(cases.sweep(quick=(V == 64))):
        if rt != routine:
            continue
        print(f"case {rt}{p} {opts} {dims} V={V} nm={nm}", flush=True)
        _check(cases.make_case(rt, p, opts, dims, V, nm, seed=20000 + i, ld_extra=i % 2))
        n += 1
    assert n > 0
