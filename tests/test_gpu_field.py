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
    # FP64: the same sums; mixed: the field kernel's FP32 near-field weights (radian form)
    # and the evaluate's (half-turn form, other lane grouping) agree to FP32 rounding
    atol = (1e-12 if precision == "fp64" else 1e-6) * np.abs(phi).max()
    np.testing.assert_allclose(phi, hf.evaluate(st, q), rtol=0, atol=atol)
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


def test_field_mixed_near_field_phase64():
    """Mixed precision at k r up to 130 between neighbour leaves (k = 300, leaves of width
    1/8): every leaf takes the FP64 coordinates and phase.  With every M2L level kept in
    FP64 the near field is the only FP32 part, so phi and E sit within the FP32 rounding of
    the 1/r weights (~1e-7) of the FP64 plan at the same degrees; an FP32 phase would err
    by about k r 2^-23 ~ 1e-5 on these pairs."""
    import hfmm0_amd as hf
    print("case: field mixed phase64 n=20000 k=300 p=14 leaf=48", flush=True)
    x, q = D.uniform_problem(20000)
    ref = hf.setup(x, 300.0, p=14, N_max=48)
    lv = list(range(2, ref.depth + 1))
    st = hf.setup(x, 300.0, p=14, N_max=48, precision="mixed", m2l_fp64_levels=lv)
    assert st.info.n_leaves_phase64 == st.info.n_leaves > 0
    phi64, E64 = hf.evaluate_field(ref, q)
    phi, E = hf.evaluate_field(st, q)
    dp, de = _rel(phi, phi64), _rel(E, E64)
    print(f"  vs fp64 plan: phi {dp:.2e} E {de:.2e}", flush=True)
    assert dp < 3e-7 and de < 3e-7
