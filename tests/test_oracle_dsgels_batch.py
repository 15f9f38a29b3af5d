"""Pin the oracle's mixed-precision uneven batch (oracle_dsgels_batch, the
library's dsgels_batch; no reference counterpart):

  - every refined problem's X against LAPACK dgels (scipy), within the
    least-squares perturbation bound, with iter >= 0 and A untouched;
  - cond 1e12 problems: iter = -31, and A and X bit-identical to the oracle's
    plain double batch gels (pinned against LAPACK and the reference in
    test_oracle_batch.py) on the same problem, X := B first;
  - shapes outside the single precision path (more than 64 rows, more than
    32 columns, underdetermined, nrhs > 8, n = 0 or nrhs = 0) and every
    transposed problem: iter = -1 and the plain double result, bit for bit;
  - the refinement itself is the one dsgels_compact uses (refine_column),
    which test_oracle_lapack.py pins against LAPACK on its own.
"""

import numpy as np
import scipy.linalg.lapack as lapack

import batchcase
import cbkt


def test_mixed_batch_against_lapack_and_plain():
    r = cbkt.rng(4100)
    shapes, ill, outside = batchcase.mixed_mix(r)
    desc, ap, bp, As, Bs = batchcase.make(r, shapes)
    for k in ill:
        m, n = shapes[k][:2]
        As[k] = batchcase.ill_conditioned(r, m, n)
        batchcase.set_a(ap, desc[k], As[k])
    apm, xpm, it, info = batchcase.oracle_mixed("N", desc, ap, bp)
    app, bpp = ap.copy(), bp.copy()
    infop = np.zeros(len(desc), np.int32)
    assert batchcase.oracle_batch("d", "N", desc, app, bpp, infop) == 0
    assert np.array_equal(info, infop)
    eps = 2.0 ** -53
    for k, d in enumerate(desc):
        m, n, q = d[:3]
        if k in outside:
            assert it[k] == -1, (k, d, it[k])
        elif k in ill:
            assert it[k] == -31, (k, d, it[k])
        if it[k] < 0:
            # the plain double gels on the same problem, X := B first
            assert np.array_equal(batchcase.get_a(apm, d), batchcase.get_a(app, d)), (k, d)
            assert np.array_equal(batchcase.get_b(xpm, d), batchcase.get_b(bpp, d)), (k, d)
            continue
        assert k not in outside and k not in ill
        assert np.array_equal(batchcase.get_a(apm, d), As[k])
        X = batchcase.get_b(xpm, d)[:n]
        _, xl, inf = lapack.dgels(As[k], Bs[k])
        assert inf == 0
        Xl = xl[:n]
        s = np.linalg.svd(As[k], compute_uv=False)
        kappa = s[0] / s[-1]
        res = np.linalg.norm(Bs[k] - As[k] @ Xl)
        bound = 20 * m * eps * (kappa + kappa ** 2 * res / (s[0] * np.linalg.norm(Xl)))
        assert np.abs(X - Xl).max() <= bound * np.abs(Xl).max(), (k, d)
    refined = [k for k in range(len(desc)) if it[k] >= 0]
    assert len(refined) >= len(desc) - len(ill) - len(outside) - 3, it


def test_mixed_batch_transposed_takes_the_double_path():
    r = cbkt.rng(4101)
    shapes = [(16, 8, 2), (20, 30, 3), (8, 8, 1)]
    desc, ap, bp, _, _ = batchcase.make(r, shapes)
    apm, xpm, it, info = batchcase.oracle_mixed("T", desc, ap, bp)
    app, bpp = ap.copy(), bp.copy()
    assert batchcase.oracle_batch("d", "T", desc, app, bpp) == 0
    assert np.all(it == -1)
    assert np.array_equal(apm, app)
    for d in desc:
        assert np.array_equal(batchcase.get_b(xpm, d), batchcase.get_b(bpp, d))
