"""Oracle (test infrastructure only): CPU restatement of the reference octree and lists.

This module is a checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product path never does.

It restates, in plain NumPy, the semantics of

* ``jaxfmm/trees/cubic.py:316-527`` (``build_fmm_tree``): Morton codes at the
  finest level, run-length boxes, bottom-up coarsening where a parent absorbs its
  children iff its count is <= N_max and the 2:1 balance check passes;
* ``jaxfmm/trees/cubic.py:530-602`` (``build_fmm_tree_auto``): automatic depth
  ``max(2, ceil(log8(N/N_max)))`` refined while the largest leaf holds more than
  ``2*N_max`` points, capped at depth 10;
* ``jaxfmm/trees/cubic.py:606-695`` (``assign_particles``): leaves in Morton order,
  points within a leaf in Morton order (ties by original index);
* ``jaxfmm/trees/lists.py:619-802`` (``_v_block``, ``_ua_block``, ``_wx_block``):
  V, adaptive U, W and X lists; ``jaxfmm/core.py:138-139`` adds leaf self pairs.

Boxes are returned in a canonical order: level ascending, Morton code ascending.
The reference's own packed order (fine-to-coarse) is an internal detail that no
result depends on.

Reference quirk kept on purpose: in the balance check (``cubic.py:466-480``) a
neighbour whose Morton code is 0 and that is absent at the parent level is looked
up as the *last box of the child level* (``searchsorted`` lands on the zero-filled
entries of the non-parent slots).  Its children adjacent to the parent must then be
leaves.  Both this oracle and the product reproduce that lookup.

Pinned by ``tests/test_oracle_tree.py``: brute-force definitions of 2:1 balance,
the leaf-cover property, and the list partition identity (every ordered pair of
leaf points is covered exactly once by self/U/W/X or by V at one ancestor level).
"""
import math
import numpy as np

NEIGH = np.array([(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)
                  if (dx, dy, dz) != (0, 0, 0)], dtype=np.int64)        # cubic.py:48-51 order


def _adj_child_mask():
    """(26, 8) bool: child c of the neighbour in direction i touches the centre box
    (``cubic.py:53-64``)."""
    mask = np.zeros((26, 8), bool)
    for i, (dx, dy, dz) in enumerate(NEIGH):
        for c in range(8):
            cx, cy, cz = c & 1, (c >> 1) & 1, (c >> 2) & 1
            ok = all(d == 0 or (d == -1 and b == 1) or (d == 1 and b == 0)
                     for d, b in ((dx, cx), (dy, cy), (dz, cz)))
            mask[i, c] = ok
    return mask


ADJ_CHILD_MASK = _adj_child_mask()


def expand_bits(v):
    v = np.asarray(v, dtype=np.uint64)
    v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
    return v


def compact_bits(v):
    v = np.asarray(v, dtype=np.uint64) & np.uint64(0x1249249249249249)
    v = (v | (v >> np.uint64(2))) & np.uint64(0x10C30C30C30C30C3)
    v = (v | (v >> np.uint64(4))) & np.uint64(0x100F00F00F00F00F)
    v = (v | (v >> np.uint64(8))) & np.uint64(0x1F0000FF0000FF)
    v = (v | (v >> np.uint64(16))) & np.uint64(0x1F00000000FFFF)
    v = (v | (v >> np.uint64(32))) & np.uint64(0x1FFFFF)
    return v


def morton3(ix, iy, iz):
    """``cubic.py:107-109``: x in bit 0, y in bit 1, z in bit 2."""
    return expand_bits(ix) | (expand_bits(iy) << np.uint64(1)) | (expand_bits(iz) << np.uint64(2))


def morton_grid(code):
    code = np.asarray(code, dtype=np.uint64)
    return (compact_bits(code).astype(np.int64), compact_bits(code >> np.uint64(1)).astype(np.int64),
            compact_bits(code >> np.uint64(2)).astype(np.int64))


def _domain(positions):
    """Tight bounding cube (``cubic.py:379-385``)."""
    pmin = positions.min(axis=0)
    pmax = positions.max(axis=0)
    return pmin, float(np.max(pmax - pmin))


def point_codes(positions, depth):
    """Morton codes of points on the 2**depth grid (``cubic.py:390-392``)."""
    pmin, extent = _domain(positions)
    grid = 2 ** depth
    scaled = (positions - pmin) / extent * grid
    g = np.clip(scaled.astype(np.uint64), 0, grid - 1)
    return morton3(g[:, 0], g[:, 1], g[:, 2])


def build_tree_fixed(positions, depth, n_max):
    """One ``build_fmm_tree`` call at a fixed depth.  Returns a dict keyed by level of
    ``(codes, counts, is_leaf)`` arrays sorted by code, plus the point codes."""
    positions = np.asarray(positions, dtype=np.float64)
    codes = point_codes(positions, depth)
    order = np.argsort(codes, kind="stable")
    uniq, counts = np.unique(codes[order], return_counts=True)
    levels = {depth: [uniq.astype(np.uint64), counts.astype(np.int64), np.ones(len(uniq), bool)]}
    for L in range(depth, 1, -1):                     # merge level L into L-1, stop at level 1
        codes_L, counts_L, leaf_L = levels[L]
        par = codes_L >> np.uint64(3)
        pcodes, inv = np.unique(par, return_inverse=True)
        pcounts = np.bincount(inv, weights=counts_L, minlength=len(pcodes)).astype(np.int64)
        count_ok = pcounts <= n_max
        if L == depth:
            balance_ok = np.ones(len(pcodes), bool)   # cubic.py:510-511: skipped at the first step
        else:
            balance_ok = _balance(pcodes, L - 1, levels, L)
        merge = count_ok & balance_ok
        keep = ~merge[inv]
        levels[L] = [codes_L[keep], counts_L[keep], leaf_L[keep]]
        levels[L - 1] = [pcodes.astype(np.uint64), pcounts, merge.copy()]
    return levels, codes, order


def _lookup(codes_sorted, query):
    """Index of ``query`` in the sorted array or -1."""
    pos = np.searchsorted(codes_sorted, query)
    pos_c = np.minimum(pos, max(len(codes_sorted) - 1, 0))
    found = (pos < len(codes_sorted)) & (codes_sorted[pos_c] == query) if len(codes_sorted) else \
        np.zeros(np.shape(query), bool)
    return np.where(found, pos_c, -1)


def _children_leaf_status(levels, Lc, parent_code):
    """(8,) status of children of ``parent_code`` at level Lc: 1 leaf, 0 internal,
    -1 absent."""
    if Lc not in levels:
        return np.full(8, -1)
    codes_c, _, leaf_c = levels[Lc]
    kids = parent_code * np.uint64(8) + np.arange(8, dtype=np.uint64)
    idx = _lookup(codes_c, kids)
    return np.where(idx >= 0, np.where(leaf_c[np.maximum(idx, 0)], 1, 0), -1)


def _balance(pcodes, Lp, levels, L):
    """2:1 balance check for parents ``pcodes`` at level Lp (``cubic.py:455-483``)."""
    gx, gy, gz = morton_grid(pcodes)
    size = 2 ** Lp
    ok = np.ones(len(pcodes), bool)
    codes_L, _, leaf_L = levels[L]
    last_code_L = codes_L[-1]                         # box ofs-1 in the reference's array
    for i, (dx, dy, dz) in enumerate(NEIGH):
        nx, ny, nz = gx + dx, gy + dy, gz + dz
        valid = (nx >= 0) & (nx < size) & (ny >= 0) & (ny < size) & (nz >= 0) & (nz < size)
        ncode = morton3(np.clip(nx, 0, size - 1), np.clip(ny, 0, size - 1), np.clip(nz, 0, size - 1))
        q = _lookup(pcodes, ncode)
        mask = ADJ_CHILD_MASK[i]
        for b in np.nonzero(valid)[0]:
            if q[b] >= 0:
                st = _children_leaf_status(levels, L, pcodes[q[b]])
            elif ncode[b] == 0:                       # reference quirk (module docstring)
                st = _children_leaf_status(levels, L + 1, last_code_L)
            else:
                continue
            if np.any(mask & (st == 0)):
                ok[b] = False
    return ok


def build_tree(positions, n_max=64, max_depth=None, depth_cap=10, refine_factor=2):
    """``build_fmm_tree_auto`` + ``assign_particles``.

    Returns a dict with canonical box arrays (level asc, Morton asc): ``level``,
    ``code``, ``count``, ``is_leaf``, ``parent`` (-1 for level 1), ``pstart`` (first
    sorted point of the box's subtree), plus ``perm`` (sorted -> original index),
    ``depth``, ``pmin``, ``extent``."""
    positions = np.asarray(positions, dtype=np.float64)
    N = len(positions)
    given = max_depth is not None
    md = int(max_depth) if given else max(2, math.ceil(math.log(max(N / max(n_max, 1), 1.0)) / math.log(8)))
    while True:
        levels, pcodes, order = build_tree_fixed(positions, md, n_max)
        maxleaf = max(int(levels[L][1][levels[L][2]].max()) for L in levels if levels[L][2].any())
        if given or maxleaf <= refine_factor * n_max or md >= depth_cap:
            break
        md += 1
    lv, cd, ct, lf = [], [], [], []
    for L in sorted(levels):
        c, n, f = levels[L]
        lv.append(np.full(len(c), L)); cd.append(c); ct.append(n); lf.append(f)
    level = np.concatenate(lv).astype(np.int64)
    code = np.concatenate(cd).astype(np.uint64)
    count = np.concatenate(ct).astype(np.int64)
    is_leaf = np.concatenate(lf)
    key = {(int(l), int(c)): i for i, (l, c) in enumerate(zip(level, code))}
    parent = np.array([key.get((int(l) - 1, int(c) >> 3), -1) if l > 1 else -1
                       for l, c in zip(level, code)], dtype=np.int64)
    # subtree start in Morton-sorted point order
    sorted_codes = pcodes[order]
    first = (code << (np.uint64(3) * (np.uint64(md) - level.astype(np.uint64))))
    pstart = np.searchsorted(sorted_codes, first).astype(np.int64)
    pmin, extent = _domain(positions)
    return dict(level=level, code=code, count=count, is_leaf=is_leaf, parent=parent,
                pstart=pstart, perm=order.astype(np.int64), depth=md, pmin=pmin, extent=extent,
                maxleaf=maxleaf, key=key)


def box_centres(tree):
    """``cubic.py:116-160``: centre = pmin + (grid + 1/2) * extent / 2**level."""
    gx, gy, gz = morton_grid(tree["code"])
    g = np.stack([gx, gy, gz], 1).astype(np.float64)
    s = tree["extent"] / (2.0 ** tree["level"])[:, None]
    return g * s + tree["extent"] * 0.5 ** (tree["level"][:, None] + 1) + tree["pmin"]


def _adj_coarse(b, c):
    return (b <= 2 * c + 2) & (2 * c <= b + 1)        # lists.py:566-568


def _adj_fine(b, f):
    return (2 * b <= f + 1) & (f <= 2 * b + 2)        # lists.py:571-573


def build_lists(tree):
    """V, U (without self), W, X as sorted (target, source) int arrays over canonical
    box indices, following ``lists.py:619-802``."""
    level, code, is_leaf, key = tree["level"], tree["code"], tree["is_leaf"], tree["key"]
    V, U, W, X = [], [], [], []

    def find(l, c):
        return key.get((int(l), int(c)), -1)

    for b in range(len(level)):
        L = int(level[b]); c = int(code[b])
        bx, by, bz = (int(v) for v in morton_grid(np.uint64(c)))
        leaf = bool(is_leaf[b])
        pc = c >> 3
        px, py, pz = bx >> 1, by >> 1, bz >> 1
        psize = 2 ** (L - 1)
        size = 2 ** L
        for dx, dy, dz in NEIGH:
            # parent colleagues (level L-1)
            nx, ny, nz = px + dx, py + dy, pz + dz
            pvalid = L - 1 >= 0 and 0 <= nx < psize and 0 <= ny < psize and 0 <= nz < psize
            if pvalid:
                ncode = int(morton3(nx, ny, nz))
                for ch in range(8):                  # V: children of parent colleagues
                    cc = ncode * 8 + ch
                    s = find(L, cc)
                    if s < 0 or cc == c:
                        continue
                    cx, cy, cz = (int(v) for v in morton_grid(np.uint64(cc)))
                    if max(abs(cx - bx), abs(cy - by), abs(cz - bz)) > 1:
                        V.append((b, s))
                if L > 1:
                    s = find(L - 1, ncode)
                    if s >= 0 and is_leaf[s]:
                        adj = _adj_coarse(bx, nx) and _adj_coarse(by, ny) and _adj_coarse(bz, nz)
                        if adj and leaf:
                            U.append((b, s))          # U (ii) coarser adjacent leaf
                        if not adj:
                            X.append((b, s))          # X: coarser non-adjacent leaf
            # colleagues (level L)
            mx, my, mz = bx + dx, by + dy, bz + dz
            if not (0 <= mx < size and 0 <= my < size and 0 <= mz < size):
                continue
            mcode = int(morton3(mx, my, mz))
            if leaf:
                s = find(L, mcode)
                if s >= 0 and is_leaf[s]:
                    U.append((b, s))                  # U (i) same-level leaf
                for ch in range(8):
                    fc = mcode * 8 + ch
                    s = find(L + 1, fc)
                    if s < 0:
                        continue
                    fx, fy, fz = (int(v) for v in morton_grid(np.uint64(fc)))
                    adj = _adj_fine(bx, fx) and _adj_fine(by, fy) and _adj_fine(bz, fz)
                    if adj and is_leaf[s]:
                        U.append((b, s))              # U (iii) finer adjacent leaf
                    if not adj:
                        W.append((b, s))              # W: finer non-adjacent box
    as_arr = lambda a: np.array(sorted(a), dtype=np.int64).reshape(-1, 2)
    return dict(V=as_arr(V), U=as_arr(U), W=as_arr(W), X=as_arr(X))
