"""The C++ host tree/list builder of libhfmm against the oracle restatement of the
reference (bit-exact: same boxes, same Morton codes, same lists, same permutation)."""
import numpy as np
import pytest

from oracle import direct as D
from oracle import tree as T


def _cases():
    rng = np.random.default_rng(21)
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


@pytest.mark.parametrize("name,x,nmax,md", list(_cases()), ids=lambda v: v if isinstance(v, str) else "")
def test_tree_matches_oracle(name, x, nmax, md):
    from hfmm0_amd.tree import build_tree
    got = build_tree(x, nmax, md)
    ref = T.build_tree(x, nmax, md)
    assert got["depth"] == ref["depth"]
    for key in ("level", "code", "count", "pstart", "parent", "is_leaf", "perm"):
        np.testing.assert_array_equal(np.asarray(got[key]).astype(np.int64), np.asarray(ref[key]).astype(np.int64),
                                      err_msg=f"{name}:{key}")
    assert got["extent"] == ref["extent"]
    if len(ref["level"]) > 4000:
        return
    lists = T.build_lists(ref)
    for which in "VUWX":
        np.testing.assert_array_equal(np.asarray(got[which], np.int64).reshape(-1, 2), lists[which],
                                      err_msg=f"{name}:{which}")


def test_validation_messages():
    from hfmm0_amd.tree import build_tree
    with pytest.raises(ValueError, match="nonempty finite"):
        build_tree(np.array([[0.0, np.nan, 1.0]]), 16)


def test_host_tree_matches_oracle_digests_300k():
    # the oracle's tree and lists of a 300k-point sphere + cluster input, as SHA-256
    # digests (tests/golden/make_tree_fixtures.py)
    import json
    import os
    import sys
    from hfmm0_amd.tree import build_tree
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    from make_tree_fixtures import digests, sphere_clusters_300k
    with open(os.path.join(here, "golden", "tree_sphere_clusters_300k.json")) as f:
        rec = json.load(f)
    x, nmax = sphere_clusters_300k()
    got = build_tree(x, nmax)
    assert got["depth"] == rec["depth"] and got["extent"] == rec["extent"]
    assert digests(got) == rec["digests"]
