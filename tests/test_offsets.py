"""The dispatcher's 32-bit offset predicate (cbk_amd/csrc/offsets.h) against
the offsets the register and team kernels actually form.

A buffer-resource kernel addresses element (i, j) of the matrix in lane l of
a wavefront at byte offset ((g - w0) * gstride + v) * sz + (i * si + j * sj)
* V * sz, with g = idx / V, v = idx % V for the lane's slot idx, and w0 the
first group of the wavefront (reg_chol_impl.h, qr_reg_impl.h, trsm_reg.hip,
ormqr_reg.hip; the team kernels span a single group). A load is in range when
offset + sz <= 2^31 - 1. Here the largest offset is found by enumerating every
lane of a wavefront and every referenced element, and the predicate must say
"fits" exactly when that enumeration does. Cases sit on both sides of the
boundary. No GPU needed: the header is host C++ compiled with g++."""

import ctypes
import os
import subprocess

import numpy as np
import pytest

import cbkt

HDR_DIR = os.path.join(cbkt.ROOT, "cbk_amd", "csrc")

SRC = r"""
#include "offsets.h"
extern "C" int fits(unsigned long long gstride, int ld, int row, int rows, int cols, int V,
                    int elem)
{
    return cbk::offsets_fit_bytes((size_t)gstride, ld, row != 0, rows, cols, V, (size_t)elem);
}
"""


@pytest.fixture(scope="module")
def fits(tmp_path_factory):
    d = tmp_path_factory.mktemp("offsets")
    src = d / "fits.cpp"
    src.write_text(SRC)
    so = d / "libfits.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-shared", "-fPIC", "-I", HDR_DIR, str(src), "-o",
                    str(so)], check=True)
    lib = ctypes.CDLL(str(so))
    lib.fits.argtypes = [ctypes.c_ulonglong] + [ctypes.c_int] * 6
    lib.fits.restype = ctypes.c_int
    return lib.fits


def brute_fits(ld, row, rows, cols, V, elem):
    """enumerate the offsets of the wavefront whose lanes span the most groups"""
    si, sj = (ld, 1) if row else (1, ld)
    gstride = ld * (rows if row else cols) * V
    w0 = 0
    lanes = np.arange(64)
    g = lanes // V
    v = lanes % V
    slot = ((g - w0) * gstride + v) * elem
    i = np.arange(max(rows, 1))[:, None]
    j = np.arange(max(cols, 1))[None, :]
    place = (i * si + j * sj) * V * elem
    last = int(slot.max()) + int(place.max())
    return gstride, last + elem <= 0x7FFFFFFF


def test_predicate_matches_enumerated_offsets(fits):
    r = np.random.default_rng(5)
    checked = boundaries = 0
    for V in (2, 4, 8, 16, 32, 64):
        for elem in (4, 8):
            for row in (False, True):
                for rows, cols in ((1, 1), (8, 8), (16, 2), (24, 24), (32, 4), (64, 32), (5, 9)):
                    # the critical leading dimension, from a coarse search
                    lo, hi = max(rows if not row else cols, 1), 1 << 30
                    while hi - lo > 1:
                        mid = (lo + hi) // 2
                        if brute_fits(mid, row, rows, cols, V, elem)[1]:
                            lo = mid
                        else:
                            hi = mid
                    lds = {lo - 1, lo, lo + 1, hi, hi + 7, max(rows, cols)}
                    lds |= {int(x) for x in r.integers(max(rows, cols), 4 * hi, 3)}
                    for ld in sorted(x for x in lds if x >= max(1, rows if not row else cols)):
                        gs, want = brute_fits(ld, row, rows, cols, V, elem)
                        got = fits(gs, ld, int(row), rows, cols, V, elem)
                        assert bool(got) == want, (V, elem, row, rows, cols, ld)
                        checked += 1
                    # the boundary itself is on both sides (where ld matters at all)
                    if not brute_fits(1 << 30, row, rows, cols, V, elem)[1]:
                        assert brute_fits(lo, row, rows, cols, V, elem)[1]
                        assert not brute_fits(hi, row, rows, cols, V, elem)[1]
                        boundaries += 1
    assert checked > 1000 and boundaries > 100
