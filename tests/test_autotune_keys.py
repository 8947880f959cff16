"""Host logic of the autotuner: problem-class keys and the wisdom file."""
import numpy as np


def test_problem_key_buckets():
    from hfmm0_amd.autotune import problem_key
    rng = np.random.default_rng(0)
    x = rng.uniform(size=(10_000, 3))
    k1 = problem_key(x, 40.0, 1e-6, "mixed")
    assert k1.startswith("v1|helmholtz|N2^13|A2^0|kD2^5|tol1e-6|mixed|")
    assert problem_key(x * 2.0, 20.0, 1e-6, "mixed") == k1          # k * extent is what counts
    assert problem_key(x, 40.0, 1e-3, "mixed") != k1
    assert problem_key(x, 40.0, 1e-6, "fp64") != k1
    assert problem_key(x * np.array([1, 1, 0.01]), 40.0, 1e-6, "mixed").split("|")[3] == "A2^7"


def test_wisdom_round_trip(tmp_path):
    from hfmm0_amd.autotune import load_wisdom, save_wisdom
    p = str(tmp_path / "w.json")
    assert load_wisdom(p) == {}
    save_wisdom({"a": {"N_max": 64, "m2l_kernel": 16, "p_level": [12, 12, 10]}}, p)
    assert load_wisdom(p)["a"]["p_level"] == [12, 12, 10]
