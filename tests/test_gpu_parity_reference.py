"""GPU parity with the reference's own evaluate on identical inputs.

phi_ref comes from the committed fixtures (oracle/kifmm.py, a restatement of the
reference's leveled KIFMM).  Parity rule: ||phi_build - phi_ref|| / ||phi_ref|| must not
exceed err_build + err_ref, both measured against the FP64 direct sum on the same inputs.
"""
import glob
import os
import sys

import numpy as np
import pytest

from oracle import direct as D

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "golden", "kifmm_*.npz"))))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_matches_reference_evaluate(path, precision):
    import hfmm0_amd as hf
    f = np.load(path)
    x, q, k, leaf = f["positions"], f["charges"], float(f["k"]), int(f["leaf"])
    print(f"case: {os.path.basename(path)} precision={precision}", flush=True)
    sys.stdout.flush()
    phi = hf.evaluate(hf.setup(x, k, p=12, N_max=leaf, precision=precision), q)
    e_b = D.relative_error(phi, f["phi_direct"])
    e_r = D.relative_error(f["phi_ref"], f["phi_direct"])
    d = D.relative_error(phi, f["phi_ref"])
    print(f"  build err {e_b:.2e} ref err {e_r:.2e} build-vs-ref {d:.2e}", flush=True)
    assert d <= e_b + e_r + 1e-12
    if k == 0:
        assert np.count_nonzero(phi.imag) == 0
