"""GPU test of the autotuner: the planned evaluator certifies the tolerance on sampled
targets, records its configuration in the wisdom file, and a second plan of the same
problem class reuses it without searching."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_plan_tunes_certifies_and_reuses_wisdom(tmp_path, monkeypatch):
    import hfmm0_amd.autotune as A
    from hfmm0_amd.learned import sampled_error
    from hfmm0_amd.workloads import mixed_problem
    monkeypatch.setattr(A, "wisdom_path", lambda: str(tmp_path / "wisdom.json"))
    print("case: autotune plan n=25000 k=15 tol=1e-6", flush=True)
    x, q, _, _ = mixed_problem(25_000, (15.0, 15.0), seed=800_001)
    fmm = A.plan(x, 15.0, tol=1e-6, verbose=True)
    phi = fmm(q)
    assert fmm.tuned
    err = sampled_error(x, q, 15.0, phi, n_samp=1024, seed=3)
    print(f"  config {fmm.config} err {err:.2e}", flush=True)
    assert err <= 2e-6
    w = A.load_wisdom()
    assert fmm.key in w and w[fmm.key]["N_max"] in (32, 64, 128)
    fmm.close()
    x2, q2, _, _ = mixed_problem(25_000, (15.0, 15.0), seed=800_002)
    fmm2 = A.plan(x2, 15.0, tol=1e-6)
    phi2 = fmm2(q2)
    assert not fmm2.tuned and fmm2.key == fmm.key or not fmm2.tuned
    assert fmm2.report[-1][1] <= 1e-6
    fmm2.close()
