"""Pins for the CPU restatement of the reference evaluate (oracle/kifmm.py).

The reference ships no golden vectors; its harness asserts (test.py:63-72) are the pins:
error against the FP64 direct sum < 5e-2 at p <= 3 and < 5e-3 at p >= 4 on the full
complex value and on both parts, linearity evaluate(-0.5 q) = -0.5 evaluate(q), an
exactly real result at k = 0, and errors that fall as p grows.  The committed fixtures
are checked for integrity and reproduced.
"""
import glob
import os

import numpy as np
import pytest

from oracle import direct as D
from oracle import kifmm

HERE = os.path.dirname(os.path.abspath(__file__))


def _problem(n):
    rng = np.random.default_rng(42 + n)
    x = rng.uniform(0.0, 1.0, (n, 3))
    return x, rng.normal(size=n)


@pytest.mark.parametrize("k", [0.0, 3.0])
def test_reference_harness_assertions(k):
    x, q = _problem(256)
    ref = D.direct_numpy(x, q, k)
    base = None
    for p in (3, 5):
        out = kifmm.evaluate(x, q, k, p=p, n_max=16)
        assert out.dtype == np.complex128 and np.isfinite(out).all()
        errs = np.array([D.relative_error(a, b) for a, b in ((out, ref), (out.real, ref.real),
                                                              (out.imag, ref.imag))])
        if k == 0:
            assert np.count_nonzero(out.imag) == 0
            errs = errs[:2]
        assert max(errs) < (5e-2 if p <= 3 else 5e-3), errs
        np.testing.assert_allclose(kifmm.evaluate(x, -0.5 * q, k, p=p, n_max=16), -0.5 * out,
                                   rtol=1e-10, atol=1e-10)
        if base is not None:
            assert np.all(errs < np.maximum(0.8 * base, 2e-6))
        base = errs


def test_surface_points():
    for p in (2, 3, 6):
        s = kifmm.surface_pts(p, 1.05)
        assert len(s) == 6 * p * p - 12 * p + 8                      # surfaces.py:24-37
        assert np.allclose(np.abs(s).max(axis=1), 1.05)
        assert len(np.unique(np.round(s, 12), axis=0)) == len(s)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "golden", "kifmm_*.npz"))))
def test_fixtures(path):
    f = np.load(path)
    x, q, k = f["positions"], f["charges"], float(f["k"])
    np.testing.assert_allclose(f["phi_direct"], D.direct_numpy(x, q, k), rtol=1e-12, atol=1e-12)
    err = D.relative_error(f["phi_ref"], f["phi_direct"])
    assert err < (5e-2 if int(f["p_ref"]) <= 3 else 5e-3)
    if len(x) <= 640:
        again = kifmm.evaluate(x, q, k, p=int(f["p_ref"]), n_max=int(f["leaf"]))
        np.testing.assert_allclose(again, f["phi_ref"], rtol=1e-11, atol=1e-11 * np.abs(f["phi_ref"]).max())


def test_c_evaluate_matches_numpy_restatement():
    # the C/OpenMP evaluate phases (oracle/kifmm_eval.c, timed by bench.py's CPU baseline)
    # against the NumPy restatement on a clustered input with W and X lists
    from oracle import kifmm as K
    from oracle import kifmm_c as KC
    rng = np.random.default_rng(77)
    n = 900
    x = rng.uniform(size=(n, 3))
    x[: n // 2] = rng.uniform(0.1, 0.3, (n // 2, 3))
    q = rng.normal(size=n) + 1j * rng.normal(size=n)
    pr = KC.Prepared(x, 3.0, p=4, n_max=12)
    assert len(pr.arrays["w_src"]) > 0 and len(pr.arrays["x_src"]) > 0
    got = pr.evaluate(q)
    ref = K.evaluate(x, q, 3.0, p=4, n_max=12)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12
    # k = 0: one pass, exactly real
    assert np.count_nonzero(KC.Prepared(x, 0.0, p=4, n_max=12).evaluate(q.real).imag) == 0
