"""Oracle (test infrastructure only): CPU restatement of the reference's HFMM evaluate.

The reference's Helmholtz path (``HFMM.py:55-91``) is a kernel-independent FMM in
"leveled" mode, run twice on the same tree: once with the kernel cos(kr)/r and once with
sin(kr)/r, returning ``real + 1j*imag``.  This module restates it in NumPy, FP64, so that
the build's results can be compared with the reference's own algorithm on identical
inputs, and so that ``bench.py`` can time the reference's algorithm on host cores.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import it.

Followed files (reference):
* surfaces ``jaxfmm/kifmm/surfaces.py:20-85`` (radii 1.05 / 2.95, p x p face grids,
  duplicate edge/corner points removed, first occurrence order);
* operators ``jaxfmm/kifmm/operators.py:72-99`` (Tikhonov-damped factored
  pseudoinverse), ``:276-425`` (per-level check/equivalent matrices, 8 M2M and L2L
  kernels, 343 M2L kernels), ``:572-641`` (one set per level at half-side
  extent * 0.5**(l+1); M2M tables of parent level pl assembled at the child scale);
* phases ``jaxfmm/kifmm/phases.py:79-280`` (P2M, M2M, M2L, L2L), ``:1533-1582`` (L2P),
  ``:1064-1140`` (M2P, W list), ``:1709-1770`` (P2L, X list);
* driver ``jaxfmm/core.py:214-335``; near field ``jaxfmm/near.py:48-312``
  (trig(kr)/r over U + self pairs, r = 0 excluded);
* kernel ``HFMM.py:12-52`` (no 1/(4 pi); the imaginary-kernel tables use eps_rel = 1e-12,
  ``HFMM.py:75-77``; k == 0 has no imaginary pass, ``HFMM.py:71-72, 90``).
Tree and lists come from :mod:`oracle.tree`.

Pinned by ``tests/test_oracle_kifmm.py``: the reference harness's own assertions
(test.py:63-72): error against the FP64 direct sum below 5e-2 / 5e-3, linearity, an
exactly real result at k = 0, errors falling as p grows.  Bitwise agreement with a JAX
run is unpinned (jax is not installed).
"""
import numpy as np

from oracle import tree as T

R_EQUIV = 1.05
R_CHECK = 2.95
EPS = np.finfo(np.float64).eps

CHILD_OFFSETS = np.array([(-1, -1, -1), (1, -1, -1), (-1, 1, -1), (1, 1, -1),
                          (-1, -1, 1), (1, -1, 1), (-1, 1, 1), (1, 1, 1)], dtype=np.float64)


def surface_pts(p, r=1.0):
    """Cube surface grid of half-side r (surfaces.py:40-85)."""
    t = np.linspace(-r, r, p)
    g0, g1 = np.meshgrid(t, t, indexing="ij")
    g0, g1 = g0.reshape(-1), g1.reshape(-1)
    c = r * np.ones(p * p)
    pts = np.vstack([np.stack([c, g0, g1], 1), np.stack([-c, g0, g1], 1),
                     np.stack([g0, c, g1], 1), np.stack([g0, -c, g1], 1),
                     np.stack([g0, g1, c], 1), np.stack([g0, g1, -c], 1)])
    key = np.round(pts / max(r, 1e-300), 14)
    _, first = np.unique(key, axis=0, return_index=True)
    return pts[np.sort(first)]


def kernel_matrix(x, y, k, trig):
    """trig(k r)/r with the reference's relative floor r > 16 eps max(r) (HFMM.py:20-27)."""
    r = np.linalg.norm(x[:, None, :] - y[None, :, :], axis=-1)
    ok = r > 16 * EPS * r.max(initial=0.0)
    return np.where(ok, trig(k * r) / np.where(ok, r, 1.0), 0.0)


def pinv_factors(A, eps_rel):
    """(U, s/(s^2+eps^2), Vt) with eps = eps_rel * s_max (operators.py:72-99)."""
    U, s, Vt = np.linalg.svd(A, full_matrices=False)
    eps = eps_rel * s[0]
    return U, s / (s * s + eps * eps), Vt


def decode(X, F):
    """X (A^+)^T applied as ((X U) * d) Vt (phases.py:21-34)."""
    U, d, Vt = F
    return ((X @ U) * d) @ Vt


def level_operators(p, s, k, trig, eps_rel):
    """One level's tables, unit geometry with the kernel scaled by the half-side s."""
    eq = surface_pts(p, R_EQUIV)
    ck = surface_pts(p, R_CHECK)
    G = lambda a, b: kernel_matrix(a * s, b * s, k, trig)
    ops = dict(F_up=pinv_factors(G(ck, eq), eps_rel), F_dn=pinv_factors(G(eq, ck), eps_rel))
    ops["M2M_K"] = np.stack([G(ck * 2.0, eq + CHILD_OFFSETS[c]) for c in range(8)])
    ops["L2L_K"] = np.stack([G(eq + CHILD_OFFSETS[c], ck * 2.0) for c in range(8)])
    m2l = {}
    for dx in range(-3, 4):
        for dy in range(-3, 4):
            for dz in range(-3, 4):
                if max(abs(dx), abs(dy), abs(dz)) <= 1:
                    continue
                m2l[(dx, dy, dz)] = G(eq + 2.0 * np.array([dx, dy, dz], float), eq)
    ops["M2L_K"] = m2l
    return ops


def _component(x, q, k, trig, p, eps_rel, tr, lists, centres, h):
    md = tr["depth"]
    level, code, is_leaf = tr["level"], tr["code"], tr["is_leaf"]
    nb = len(level)
    eq_pts = surface_pts(p, R_EQUIV)
    ck_pts = surface_pts(p, R_CHECK)
    ns = len(eq_pts)
    lv_ops = [level_operators(p, tr["extent"] * 0.5 ** (l + 1), k, trig, eps_rel) for l in range(md + 1)]
    perm, pstart, count = tr["perm"], tr["pstart"], tr["count"]
    pts = lambda b: x[perm[pstart[b]:pstart[b] + count[b]]]
    chg = lambda b: q[perm[pstart[b]:pstart[b] + count[b]]]
    idx = lambda b: perm[pstart[b]:pstart[b] + count[b]]
    g = lambda a, b: kernel_matrix(a, b, k, trig)
    # P2M (phases.py:79-125): physical check surface, the leaf level's decode
    up = np.zeros((nb, ns))
    for b in np.nonzero(is_leaf)[0]:
        check = ck_pts * h[b] + centres[b]
        up[b] = decode((g(check, pts(b)) @ chg(b))[None, :], lv_ops[level[b]]["F_up"])[0]
    # M2M (phases.py:128-173): child level md..2 into parent level md-1..1
    for cl in range(md, 1, -1):
        pl = cl - 1
        check = np.zeros((nb, ns))
        for c in np.nonzero(level == cl)[0]:
            check[tr["parent"][c]] += lv_ops[cl]["M2M_K"][int(code[c]) & 7] @ up[c]
        par = np.nonzero(level == pl)[0]
        up[par] += decode(check[par], lv_ops[pl]["F_up"])
    # M2L (phases.py:176-237): raw kernels into check potentials, one decode per box
    check = np.zeros((nb, ns))
    groups = {}
    for t, s in lists["V"]:
        gt = T.morton_grid(code[t]); gs = T.morton_grid(code[s])
        key = (int(level[t]), int(gt[0] - gs[0]), int(gt[1] - gs[1]), int(gt[2] - gs[2]))
        groups.setdefault(key, []).append((t, s))
    for (l, dx, dy, dz), pairs in groups.items():
        pr = np.array(pairs)
        np.add.at(check, pr[:, 0], up[pr[:, 1]] @ lv_ops[l]["M2L_K"][(dx, dy, dz)].T)
    dn = np.zeros((nb, ns))
    for l in range(md + 1):
        sel = np.nonzero(level == l)[0]
        if len(sel):
            dn[sel] = decode(check[sel], lv_ops[l]["F_dn"])
    # P2L, X list (phases.py:1709-1770): downward check surface = R_EQUIV
    for t, s in lists["X"]:
        chk = eq_pts * h[t] + centres[t]
        dn[t] += decode((g(chk, pts(s)) @ chg(s))[None, :], lv_ops[level[t]]["F_dn"])[0]
    # L2L (phases.py:240-280): child levels 2..md in order
    for cl in range(2, md + 1):
        for c in np.nonzero(level == cl)[0]:
            par = tr["parent"][c]
            dn[c] += decode((lv_ops[cl]["L2L_K"][int(code[c]) & 7] @ dn[par])[None, :], lv_ops[cl]["F_dn"])[0]
    pot = np.zeros(len(x))
    # L2P (phases.py:1533-1582): downward equivalent surface = R_CHECK
    for b in np.nonzero(is_leaf)[0]:
        pot[idx(b)] += g(pts(b), ck_pts * h[b] + centres[b]) @ dn[b]
    # M2P, W list (phases.py:1064-1140): upward equivalent surface = R_EQUIV
    for t, s in lists["W"]:
        pot[idx(t)] += g(pts(t), eq_pts * h[s] + centres[s]) @ up[s]
    # near field (near.py:48-312): U + self pairs, r = 0 excluded
    nbr = {}
    for t, s in lists["U"]:
        nbr.setdefault(int(t), []).append(int(s))
    for t in np.nonzero(is_leaf)[0]:
        srcs = [int(t)] + nbr.get(int(t), [])
        si = np.concatenate([idx(s) for s in srcs])
        r = np.linalg.norm(pts(t)[:, None, :] - x[si][None, :, :], axis=-1)
        w = np.where(r > 0, trig(k * r) / np.where(r > 0, r, 1.0), 0.0)
        pot[idx(t)] += w @ q[si]
    return pot


def evaluate(positions, charges, k, p=4, n_max=16, max_depth=None):
    """``HFMM.setup`` + ``HFMM.evaluate`` for real charges (reference semantics);
    complex charges are handled by linearity, as the build's parity rule states."""
    x = np.asarray(positions, dtype=np.float64)
    q = np.asarray(charges)
    if np.iscomplexobj(q):
        return (evaluate(x, q.real, k, p, n_max, max_depth)
                + 1j * evaluate(x, q.imag, k, p, n_max, max_depth))
    q = q.astype(np.float64)
    tr = T.build_tree(x, n_max, max_depth)
    lists = T.build_lists(tr)
    centres = T.box_centres(tr)
    h = tr["extent"] * 0.5 ** (tr["level"].astype(np.float64) + 1.0)
    real = _component(x, q, float(k), np.cos, p, 1e-7, tr, lists, centres, h)
    if float(k) == 0.0:
        return real + 1j * np.zeros_like(real)
    imag = _component(x, q, float(k), np.sin, p, 1e-12, tr, lists, centres, h)
    return real + 1j * imag
