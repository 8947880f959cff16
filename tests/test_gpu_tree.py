"""GPU tests of the device tree and list builder (rocPRIM sort / run-length / reduce-by-key
/ select plus the project's merge and list kernels) against the oracle restatement of the
reference's builders (oracle/tree.py: jaxfmm/trees/cubic.py:316-695, lists.py:563-931):

* element for element equal to the oracle's tree on uniform, clustered, surface, film and
  degenerate inputs, and to its V/U/W/X lists wherever the oracle's list builder runs in
  seconds (up to 5000 boxes);
* on a 300k-point sphere + cluster input, equal to the SHA-256 digests of the oracle's
  tree and lists (tests/golden/tree_sphere_clusters_300k.json, written by
  tests/golden/make_tree_fixtures.py);
* a single point: the build's documented deviation (extent 1 where the reference divides
  by a zero extent), compared with the host builder.

A plan built on the device tree evaluates identically to one built on the host tree."""
import json
import os

import numpy as np
import pytest

from oracle import tree as T

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TREE_KEYS = ("level", "code", "count", "pstart", "parent", "is_leaf", "perm")


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
    yield "duplicates", np.repeat(rng.uniform(size=(40, 3)), 3, axis=0), 4, None
    yield "c1_20k", D.uniform_problem(20000)[0], 64, None
    x = rng.uniform(size=(3000, 3)) * np.array([1.0, 0.3, 0.01])
    yield "thin_film", x, 16, None
    c = rng.normal(size=(30_000, 3))
    yield "sphere_30k", 0.5 + 0.45 * c / np.linalg.norm(c, axis=1, keepdims=True), 32, None


@pytest.mark.parametrize("name,x,nmax,md", list(_cases()), ids=lambda v: v if isinstance(v, str) else "")
def test_device_tree_equals_oracle(name, x, nmax, md):
    from hfmm0_amd.tree import build_tree
    print(f"case: device tree vs oracle {name} n={len(x)} n_max={nmax} max_depth={md}", flush=True)
    got = build_tree(x, nmax, md, device=True)
    ref = T.build_tree(x, nmax, md)
    assert got["depth"] == ref["depth"] and got["maxleaf"] == ref["maxleaf"]
    assert got["extent"] == ref["extent"]
    for key in TREE_KEYS:
        np.testing.assert_array_equal(np.asarray(got[key]).astype(np.int64), np.asarray(ref[key]).astype(np.int64),
                                      err_msg=f"{name}:{key}")
    if len(ref["level"]) <= 5000:
        lists = T.build_lists(ref)
        for which in "VUWX":
            np.testing.assert_array_equal(np.asarray(got[which], np.int64).reshape(-1, 2), lists[which],
                                          err_msg=f"{name}:{which}")


def test_device_tree_matches_oracle_digests_300k():
    import sys
    from hfmm0_amd.tree import build_tree
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_tree_fixtures import digests, sphere_clusters_300k
    with open(os.path.join(HERE, "golden", "tree_sphere_clusters_300k.json")) as f:
        rec = json.load(f)
    x, nmax = sphere_clusters_300k()
    print(f"case: device tree vs oracle digests sphere_clusters_300k n={len(x)} n_max={nmax}", flush=True)
    got = build_tree(x, nmax, None, device=True)
    assert got["depth"] == rec["depth"] and got["extent"] == rec["extent"]
    assert digests(got) == rec["digests"]


def test_single_point_deviation_matches_host():
    from hfmm0_amd.tree import build_tree
    x = np.random.default_rng(31).uniform(size=(1, 3))
    print("case: device tree single point", flush=True)
    got = build_tree(x, 16, None, device=True)
    ref = build_tree(x, 16, None)
    assert got["extent"] == ref["extent"] == 1.0
    for key in TREE_KEYS + tuple("VUWX"):
        np.testing.assert_array_equal(got[key], ref[key], err_msg=key)


def test_plan_on_device_tree_evaluates_identically():
    import hfmm0_amd as hf
    from oracle import direct as D
    print("case: plan with device tree n=30000 k=10 p=10", flush=True)
    x, q = D.uniform_problem(30000)
    a = hf.evaluate(hf.setup(x, 10.0, p=10, N_max=32, tree_device=True), q)
    b = hf.evaluate(hf.setup(x, 10.0, p=10, N_max=32), q)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * np.abs(b).max())
