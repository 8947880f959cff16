"""GPU tests of the field E = -grad phi: against the oracle's FP64 direct field on
sampled targets (uniform and adaptive trees, FP64 and mixed precision, k = 0 real
charges and negative k), with phi equal to the plain evaluate."""
import numpy as np
import pytest

from oracle import direct as D

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("k,p,precision,tol", [(10.0, 12, "fp64", 1e-5), (10.0, 12, "mixed", 3e-5),
                                               (0.0, 10, "fp64", 3e-5), (-6.0, 12, "fp64", 1e-5)])
def test_field_uniform(k, p, precision, tol):
    import hfmm0_amd as hf
    print(f"case: field uniform n=20000 k={k} p={p} {precision}", flush=True)
    x, q = D.uniform_problem(20000)
    if k == 0.0:
        q = q.real.copy()
    st = hf.setup(x, k, p=p, N_max=48, precision=precision)
    phi, E = hf.evaluate_field(st, q)
    idx = D.sample_targets(20000, 300)
    ref = D.direct_field_c(x, q, k, idx)
    e = _rel(E[idx], ref)
    print(f"  field rel err {e:.2e}", flush=True)
    assert e < tol
    np.testing.assert_allclose(phi, hf.evaluate(st, q), rtol=0, atol=1e-12 * np.abs(phi).max())
    if k == 0.0:
        assert np.all(E.imag == 0)


def test_field_adaptive_clustered():
    import hfmm0_amd as hf
    print("case: field clustered n=15000 k=8 p=12 (W and X lists)", flush=True)
    rng = np.random.default_rng(5)
    x = np.vstack([rng.uniform(size=(5000, 3)), rng.normal(0.3, 0.02, size=(10000, 3))])
    q = rng.normal(size=15000) + 1j * rng.normal(size=15000)
    st = hf.setup(x, 8.0, p=12, N_max=24)
    t = st.tree()
    assert len(t["W"]) and len(t["X"])
    phi, E = hf.evaluate_field(st, q)
    idx = D.sample_targets(15000, 300)
    e = _rel(E[idx], D.direct_field_c(x, q, 8.0, idx))
    print(f"  field rel err {e:.2e}", flush=True)
    assert e < 5e-6


def test_field_converges_with_degree():
    import hfmm0_amd as hf
    print("case: field convergence n=20000 k=10 p=8 -> 16", flush=True)
    x, q = D.uniform_problem(20000)
    idx = D.sample_targets(20000, 300)
    ref = D.direct_field_c(x, q, 10.0, idx)
    errs = []
    for p in (8, 12, 16):
        _, E = hf.evaluate_field(hf.setup(x, 10.0, p=p, N_max=48), q)
        errs.append(_rel(E[idx], ref))
    print(f"  field errors {errs}", flush=True)
    assert errs[1] < errs[0] / 5 and errs[2] < errs[1] / 5 and errs[2] < 5e-7
