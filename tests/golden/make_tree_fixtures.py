"""Writes tests/golden/tree_sphere_clusters_300k.json: SHA-256 digests of the oracle's
octree and lists (oracle/tree.py, the restatement of jaxfmm/trees/cubic.py:316-695 and
jaxfmm/trees/lists.py:563-931) for a 300k-point sphere + Gaussian-cluster input that is
too large to rebuild with the oracle inside a test (its lists take about two minutes on
one core).  The GPU and host builders are checked against these digests.

Run:  python tests/golden/make_tree_fixtures.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

KEYS = ("level", "code", "count", "pstart", "parent", "is_leaf", "perm", "V", "U", "W", "X")


def sphere_clusters_300k():
    rng = np.random.default_rng(3_000_031)
    s = rng.normal(size=(150_000, 3))
    g = rng.normal(size=(150_000, 3)) * 0.02 + rng.uniform(0.1, 0.9, (8, 3))[rng.integers(0, 8, 150_000)]
    return np.vstack([0.5 + 0.45 * s / np.linalg.norm(s, axis=1, keepdims=True), g]), 32


def digests(arrays):
    """SHA-256 of each array as little-endian int64 (lists as (n, 2) rows)."""
    out = {}
    for k in KEYS:
        a = np.ascontiguousarray(np.asarray(arrays[k]).astype("<i8").reshape(-1, 2 if k in "VUWX" else 1))
        out[k] = {"sha256": hashlib.sha256(a.tobytes()).hexdigest(), "n": int(a.shape[0])}
    return out


def main():
    from oracle import tree as T
    x, nmax = sphere_clusters_300k()
    t = T.build_tree(x, nmax)
    lists = T.build_lists(t)
    arrays = dict(t)
    arrays.update(lists)
    rec = {"input": "sphere_clusters_300k (default_rng(3_000_031)), n_max 32", "depth": int(t["depth"]),
           "extent": float(t["extent"]), "digests": digests(arrays)}
    with open(os.path.join(HERE, "tree_sphere_clusters_300k.json"), "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps(rec, indent=1))


if __name__ == "__main__":
    main()
