"""GPU tests of the device tree and list builder: element for element equal to the host
builder (itself pinned against the oracle restatement of the reference) on uniform,
clustered, surface and degenerate inputs, and a plan built on it evaluates identically."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cases():
    from oracle import direct as D
    rng = np.random.default_rng(31)
    yield "uniform3000", rng.uniform(size=(3000, 3)), 16, None
    x = rng.uniform(size=(2000, 3))
    x[:1000] = rng.uniform(0.1, 0.25, (1000, 3))
    yield "clustered", x, 16, None
    c = rng.normal(size=(2500, 3))
    yield "sphere", 0.5 + 0.45 * c / np.linalg.norm(c, axis=1, keepdims=True), 8, None
    g = rng.normal(size=(1200, 3)) * 0.02 + rng.uniform(0.1, 0.9, (8, 3))[rng.integers(0, 8, 1200)]
    yield "gaussians", g, 12, None
    yield "fixed_depth", rng.uniform(size=(500, 3)), 16, 4
    yield "tiny", rng.uniform(size=(5, 3)), 16, None
    yield "single", rng.uniform(size=(1, 3)), 16, None
    yield "duplicates", np.repeat(rng.uniform(size=(40, 3)), 3, axis=0), 4, None
    yield "c1_20k", D.uniform_problem(20000)[0], 64, None
    x = rng.uniform(size=(3000, 3)) * np.array([1.0, 0.3, 0.01])
    yield "thin_film", x, 16, None
    s = rng.normal(size=(150_000, 3))
    x = np.vstack([0.5 + 0.45 * s / np.linalg.norm(s, axis=1, keepdims=True),
                   rng.normal(size=(150_000, 3)) * 0.02 + rng.uniform(0.1, 0.9, (8, 3))[rng.integers(0, 8, 150_000)]])
    yield "sphere_clusters_300k", x, 32, None


@pytest.mark.parametrize("name,x,nmax,md", list(_cases()), ids=lambda v: v if isinstance(v, str) else "")
def test_device_tree_equals_host_tree(name, x, nmax, md):
    from hfmm0_amd.tree import build_tree
    print(f"case: device tree {name} n={len(x)} n_max={nmax} max_depth={md}", flush=True)
    got = build_tree(x, nmax, md, device=True)
    ref = build_tree(x, nmax, md)
    assert got["depth"] == ref["depth"] and got["maxleaf"] == ref["maxleaf"]
    assert got["extent"] == ref["extent"]
    for key in ("level", "code", "count", "pstart", "parent", "is_leaf", "perm", "V", "U", "W", "X"):
        np.testing.assert_array_equal(got[key], ref[key], err_msg=f"{name}:{key}")


def test_plan_on_device_tree_evaluates_identically():
    import hfmm0_amd as hf
    from oracle import direct as D
    print("case: plan with device tree n=30000 k=10 p=10", flush=True)
    x, q = D.uniform_problem(30000)
    a = hf.evaluate(hf.setup(x, 10.0, p=10, N_max=32, tree_device=True), q)
    b = hf.evaluate(hf.setup(x, 10.0, p=10, N_max=32), q)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * np.abs(b).max())
