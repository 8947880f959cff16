"""Pins for the octree/list oracle against brute-force definitions.

* every point lies geometrically in exactly one leaf, with leaf counts matching;
* a leaf above the finest level holds at most N_max points (it was created by a merge);
* 2:1 balance: touching leaves differ by at most one level (brute force over pairs);
* list partition identity: every ordered pair of leaves (A <- B) is covered exactly
  once, by self/U (direct), by W (A, ancestor-or-self of B), by X (ancestor-or-self
  of A, B), or by V at one level (ancestors of both);
* geometry of each list: V/X/W pairs are non-adjacent, U pairs touch.
"""
import numpy as np
import pytest

from oracle import tree as T


def _problems():
    rng = np.random.default_rng(11)
    yield "uniform", rng.uniform(size=(3000, 3)), 16, None
    x = rng.uniform(size=(2000, 3))
    x[:1000] = rng.uniform(0.1, 0.25, (1000, 3))                  # test.py --clustered
    yield "clustered", x, 16, None
    c = rng.normal(size=(1500, 3))
    x = 0.5 + 0.45 * c / np.linalg.norm(c, axis=1, keepdims=True)  # sphere surface
    yield "sphere", x, 8, None
    yield "fixed_depth", rng.uniform(size=(700, 3)), 16, 3


def _boxes_touch(tr, a, b):
    """Closed boxes a and b intersect (share a face, edge or corner) and are disjoint
    in interior."""
    la, lb = tr["level"][a], tr["level"][b]
    L = max(la, lb)
    ga = np.array(T.morton_grid(tr["code"][a]), dtype=np.int64) << (L - la)
    gb = np.array(T.morton_grid(tr["code"][b]), dtype=np.int64) << (L - lb)
    sa, sb = 1 << (L - la), 1 << (L - lb)
    lo = np.maximum(ga, gb)
    hi = np.minimum(ga + sa, gb + sb)
    return bool(np.all(lo <= hi)) and not bool(np.all(lo < hi))


def _ancestors(tr, b):
    out = [b]
    while tr["parent"][out[-1]] >= 0:
        out.append(int(tr["parent"][out[-1]]))
    return out


@pytest.mark.parametrize("name,x,nmax,md", list(_problems()), ids=lambda v: v if isinstance(v, str) else "")
def test_tree_and_lists(name, x, nmax, md):
    tr = T.build_tree(x, nmax, md)
    leaves = np.nonzero(tr["is_leaf"])[0]
    # --- point cover
    pts_in = np.zeros(len(x), int)
    depth = tr["depth"]
    pcodes = T.point_codes(x, depth)
    for b in leaves:
        L = tr["level"][b]
        s, n = tr["pstart"][b], tr["count"][b]
        orig = tr["perm"][s:s + n]
        pts_in[orig] += 1
        assert np.all((pcodes[orig] >> np.uint64(3 * (depth - L))) == tr["code"][b])
        if L < depth:
            assert n <= nmax
    assert np.all(pts_in == 1)
    # internal boxes: count == sum of children
    for b in np.nonzero(~tr["is_leaf"])[0]:
        kids = np.nonzero(tr["parent"] == b)[0]
        assert len(kids) > 0 and tr["count"][kids].sum() == tr["count"][b]
    # --- 2:1 balance
    for i, a in enumerate(leaves):
        for b in leaves[i + 1:]:
            if abs(int(tr["level"][a]) - int(tr["level"][b])) > 1:
                assert not _boxes_touch(tr, a, b), (a, b)
    # --- lists
    lists = T.build_lists(tr)
    for a, b in lists["U"]:
        assert tr["is_leaf"][a] and tr["is_leaf"][b] and _boxes_touch(tr, a, b)
    for a, b in lists["V"]:
        assert tr["level"][a] == tr["level"][b] and not _boxes_touch(tr, a, b)
        assert _boxes_touch(tr, tr["parent"][a], tr["parent"][b]) or tr["parent"][a] == tr["parent"][b]
    for a, b in lists["W"]:
        assert tr["is_leaf"][a] and tr["level"][b] == tr["level"][a] + 1 and not _boxes_touch(tr, a, b)
    for a, b in lists["X"]:
        assert tr["is_leaf"][b] and tr["level"][b] == tr["level"][a] - 1 and not _boxes_touch(tr, a, b)
    # W and X are dual lists
    assert set(map(tuple, lists["W"])) == set((b, a) for a, b in lists["X"])
    # --- partition identity over ordered leaf pairs
    anc = {int(b): _ancestors(tr, int(b)) for b in leaves}
    U = set(map(tuple, lists["U"])) | set((int(b), int(b)) for b in leaves)
    V = set(map(tuple, lists["V"]))
    W = set(map(tuple, lists["W"]))
    X = set(map(tuple, lists["X"]))
    for a in leaves:
        for b in leaves:
            a, b = int(a), int(b)
            cover = int((a, b) in U)
            cover += sum((a, s) in W for s in anc[b])
            cover += sum((t, b) in X for t in anc[a])
            cover += sum((t, s) in V for t in anc[a] for s in anc[b] if tr["level"][t] == tr["level"][s])
            assert cover == 1, (name, a, b, cover)


def test_uniform_20k_leaf64_counts():
    """Uniform cube N=20k, leaf 64: depth 3, 512 leaves, 56,448 V pairs, 10,648 U+self pairs."""
    from oracle import direct
    x, _ = direct.uniform_problem(20000)
    tr = T.build_tree(x, 64)
    assert tr["depth"] == 3 and int(tr["is_leaf"].sum()) == 512
    lists = T.build_lists(tr)
    assert len(lists["V"]) == 56448
    assert len(lists["U"]) + 512 == 10648
    assert len(lists["W"]) == 0 and len(lists["X"]) == 0
