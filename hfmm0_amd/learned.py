"""Learned truncation degree: per-level features, the MLP (HIP kernels via the C ABI),
label generation by the FMM itself, and the certified evaluate with the plain fallback.

Role in the reference: order selection belongs to ``jaxfmm/tune.py``'s
``AutoEvaluator`` (certify on sampled targets, climb a (p, theta) ladder,
``tune.py:375-516``).  Here a small network predicts the degree of every tree level
from cheap features; the plain fixed-degree FMM is both its fallback and its check.

Features of level l (normalised to O(1)):
  0 level / 10
  1 k * box diameter / 50
  2 log2(mean points per box of the level) / 10      (source count)
  3 log2(max points per box of the level) / 10       (target count)
  4 log10(||q||_2 / sqrt(N))                          (charge norm)
  5 the level's share of all V-list pairs            (neighbour count)
Target: (p_min - 6) / 26, p_min the smallest degree whose truncation error at that
level, with every other level at the table degree, stays below the level budget.  The
budget of a level is the total budget times its share of the V pairs, so the error is
spent where the work is (the finest levels) and the coarse, cheap levels get high
degrees -- the allocation that minimises the M2L work for a total error.
"""
import ctypes
import math
import os

import numpy as np

from . import _lib as L

DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "degree_mlp.npz")
N_IN = 6
N_PARAMS = 8833
P_LO, P_SPAN = 6, 26


def init_params(seed):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_out, fan_in, bias in ((64, 6, 64), (64, 64, 64), (64, 64, 64), (1, 64, 1)):
        a = 1.0 / math.sqrt(fan_in)
        parts.append(rng.uniform(-a, a, fan_out * fan_in).astype(np.float32))
        parts.append(np.zeros(bias, np.float32))
    out = np.concatenate(parts)
    assert out.size == N_PARAMS
    return out


class DegreeMLP:
    """The network on the GPU (BF16-rounded weights, FP32 accumulation)."""

    def __init__(self, params=None, seed=0):
        p = init_params(seed) if params is None else np.ascontiguousarray(params, dtype=np.float32)
        if p.size != N_PARAMS:
            raise ValueError(f"expected {N_PARAMS} parameters")
        h = ctypes.c_void_p()
        L.check(L.lib().hfmm_mlp_create(L.ptr(p), ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h is not None:
            L.lib().hfmm_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        out = np.empty(N_PARAMS, np.float32)
        L.check(L.lib().hfmm_mlp_params(self._h, L.ptr(out)))
        return out

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, N_IN)
        y = np.empty(len(X), np.float32)
        L.check(L.lib().hfmm_mlp_forward(self._h, L.ptr(X), len(X), L.ptr(y)))
        return y

    def gradient(self, X, y):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, N_IN)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.empty(N_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        L.check(L.lib().hfmm_mlp_gradient(self._h, L.ptr(X), L.ptr(y), len(X), L.ptr(g), L.ptr(loss)))
        return float(loss[0]), g

    def train(self, X, y, epochs=200, batch=64, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, N_IN)
        y = np.ascontiguousarray(y, dtype=np.float32)
        hist = np.empty(epochs, np.float32)
        L.check(L.lib().hfmm_mlp_train(self._h, L.ptr(X), L.ptr(y), len(X), int(epochs), int(batch),
                                       ctypes.c_float(lr), int(seed), L.ptr(hist)))
        return hist

    def predict_degrees(self, feats, margin=1.0, p_min=4, p_max=23):
        """Per-level degrees from features: ceil(prediction + margin), clamped."""
        raw = self.forward(feats) * P_SPAN + P_LO
        return np.clip(np.ceil(raw + margin), p_min, p_max).astype(int)

    def save(self, path):
        np.savez_compressed(path, params=self.params)

    @classmethod
    def load(cls, path=None):
        """The network saved at `path` (default: the weights shipped with the package)."""
        path = DEFAULT_WEIGHTS if path is None else path
        if not os.path.exists(path):
            raise FileNotFoundError(f"no trained degree network at {path} "
                                    "(python -m hfmm0_amd.train_degrees writes one)")
        with np.load(path, allow_pickle=False) as f:
            return cls(params=f["params"])


P_CAP = 40   # largest degree of the evaluator


def degree_floor(k, extent, level):
    """High-frequency floor on the degree of a level for a 1e-6 target: ceil(0.75 kD + 2),
    D = 2 sqrt(3) h the box diameter (h = extent 2^-(level+1)), capped at P_CAP.  Measured
    on 400k points on a sphere and in Gaussian clusters at k = 100 (kD = 43 at level 2):
    level-2 degrees 31 / 28 / 26 gave 6.9e-7 / 2.0e-6 / 1.2e-5 with the finer levels as
    predicted.  The network was trained on k <= 60 with degrees <= 23, below which the
    floor does not act at the usual leaf sizes."""
    kd = abs(k) * 2.0 * math.sqrt(3.0) * extent * 0.5 ** (level + 1)
    return min(P_CAP, int(math.ceil(0.75 * kd + 2.0)))


def predict_p_level(net, positions, charges, k, n_max=16, max_depth=None, margin=1.0, p_min=4, p_max=23):
    """Per-level degrees for ``setup(p_level=...)`` (index 0 = level 1) predicted from the
    host tree's level features, raised to the high-frequency floor (degree_floor) where
    k times the box diameter is large.  Level 1 has no far field; it takes the level-2
    degree."""
    from .tree import build_tree
    tree = build_tree(positions, n_max, max_depth)
    if tree["depth"] < 2:
        return [p_min] * max(tree["depth"], 1)
    d = net.predict_degrees(level_features(tree, charges, k), margin=margin, p_min=p_min, p_max=p_max)
    d = [max(int(v), degree_floor(k, tree["extent"], lv)) for lv, v in enumerate(d, start=2)]
    return [d[0]] + d


def m2l_flops(p_level, v_per_level):
    """Algorithmic M2L flops of one evaluate: F_M2L(p_l) per V pair of level l."""
    from .workloads import f_m2l
    return sum(f_m2l(p_level[lv - 1]) * n for lv, n in v_per_level.items() if lv >= 2)


def level_features(tree, charges, k):
    """Feature rows for levels 2..depth of a tree dict (hfmm0_amd.tree.build_tree)."""
    q = np.asarray(charges)
    n = len(q)
    qn = math.log10(max(float(np.linalg.norm(q)) / math.sqrt(n), 1e-30))
    level, count = np.asarray(tree["level"]), np.asarray(tree["count"])
    V = np.asarray(tree["V"])
    vt_level = level[V[:, 0]] if len(V) else np.zeros(0, int)
    n_v = max(len(V), 1)
    rows = []
    for lv in range(2, tree["depth"] + 1):
        sel = level == lv
        nb = int(sel.sum())
        if nb == 0:
            rows.append([lv / 10.0, 0.0, 0.0, 0.0, qn, 0.0])
            continue
        diam = 2.0 * math.sqrt(3.0) * tree["extent"] * 0.5 ** (lv + 1)
        mean_c = float(count[sel].mean())
        max_c = float(count[sel].max())
        share = float((vt_level == lv).sum()) / n_v
        rows.append([lv / 10.0, k * diam / 50.0, math.log2(max(mean_c, 1.0)) / 10.0,
                     math.log2(max(max_c, 1.0)) / 10.0, qn, share])
    return np.array(rows, dtype=np.float32)


def v_shares(tree):
    """Share of all V-list pairs per level 2..depth."""
    level = np.asarray(tree["level"])
    V = np.asarray(tree["V"])
    vt = level[V[:, 0]] if len(V) else np.zeros(0, int)
    return np.array([float((vt == lv).sum()) / max(len(V), 1) for lv in range(2, tree["depth"] + 1)])


def label_levels(state, q, budget=5e-7, p_lo=4, shares=None):
    """Smallest degree per level (2..depth) keeping that level's truncation error, with
    the other levels at their table degree, at or below its budget (relative to the
    result at the table degrees): `budget` times the level's V-pair share when `shares`
    is given, else `budget`.  Uses evaluate-time degrees: no new setup per trial."""
    from .api import evaluate
    table = state.p_table
    state.set_degrees([0] * state.depth)
    ref = evaluate(state, q)
    nref = np.linalg.norm(ref)
    labels = []
    for lv in range(2, state.depth + 1):
        b = budget if shares is None else budget * max(float(shares[lv - 2]), 1e-4)
        lo, hi = p_lo, table[lv - 1]
        while lo < hi:                      # smallest p with error <= b
            mid = (lo + hi) // 2
            degs = [0] * state.depth
            degs[lv - 1] = mid
            state.set_degrees(degs)
            err = np.linalg.norm(evaluate(state, q) - ref) / nref
            if err <= b:
                hi = mid
            else:
                lo = mid + 1
        labels.append(lo)
    state.set_degrees([0] * state.depth)
    return np.array(labels)


DEFAULT_COSTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "m2l_cost.json")


def m2l_pair_cost(path=None):
    """Cost of one M2L pair as a function of the degree: the measured nanoseconds per V
    pair of the matrix-core kernel (hfmm0_amd.train_degrees --calibrate writes the table)
    where present; beyond the table, and without it, F_M2L(p) scaled to the table's last
    entry (or the flop count itself)."""
    import json
    from .workloads import f_m2l
    try:
        with open(path or DEFAULT_COSTS) as f:
            tab = {int(p): float(v) for p, v in json.load(f)["ns_per_pair"].items()}
    except (OSError, ValueError, KeyError):
        return f_m2l
    top = max(tab)

    def cost(p):
        if p in tab:
            return tab[p]
        if p > top:
            return tab[top] * f_m2l(p) / f_m2l(top)
        return tab[min(tab)] * f_m2l(p) / f_m2l(min(tab))
    return cost


def label_levels_joint(state, q, budget=6e-7, p_lo=4, cost=None):
    """Cost-optimal degrees per level (2..depth) for a total error budget: starting from
    the table degrees, lower one level's degree at a time -- the level whose step saves
    the most M2L cost (its V pairs times the drop of the per-pair cost) per unit of added
    error -- while the error of the whole evaluate (relative to the result at the table
    degrees) stays within `budget`.  `cost(p)`: per-pair cost, by default the measured
    time (m2l_pair_cost).  Every trial is one evaluate with evaluate-time degrees."""
    from .api import evaluate
    f_m2l = cost or m2l_pair_cost()
    table = state.p_table
    depth = state.depth
    tree = state.tree()
    level = np.asarray(tree["level"])
    V = np.asarray(tree["V"])
    vt = level[V[:, 0]] if len(V) else np.zeros(0, int)
    pairs = {lv: int((vt == lv).sum()) for lv in range(2, depth + 1)}
    state.set_degrees([0] * depth)
    ref = evaluate(state, q)
    nref = np.linalg.norm(ref)
    cur = {lv: table[lv - 1] for lv in range(2, depth + 1)}

    def err_of(degs):
        state.set_degrees([degs.get(lv, table[lv - 1]) for lv in range(1, depth + 1)])
        return float(np.linalg.norm(evaluate(state, q) - ref) / nref)

    e_cur = 0.0
    while True:
        best = None
        for lv in cur:
            if cur[lv] <= p_lo or pairs[lv] == 0:
                continue
            trial = dict(cur)
            trial[lv] -= 1
            e = err_of(trial)
            if e > budget:
                continue
            gain = pairs[lv] * max(f_m2l(cur[lv]) - f_m2l(cur[lv] - 1), 0.0) / max(e - e_cur, 1e-15)
            if best is None or gain > best[0]:
                best = (gain, lv, e)
        if best is None:
            break
        cur[best[1]] -= 1
        e_cur = best[2]
    state.set_degrees([0] * depth)
    return np.array([cur[lv] for lv in range(2, depth + 1)])


def sampled_error(positions, charges, k, phi, n_samp=256, seed=0):
    """Relative error of phi against the GPU FP64 direct sum on sampled targets."""
    from .api import direct
    n = len(positions)
    idx = np.sort(np.random.default_rng(seed).choice(n, min(n_samp, n), replace=False))
    ref = direct(positions, charges, k, idx)
    return float(np.linalg.norm(phi[idx] - ref) / max(np.linalg.norm(ref), 1e-300))


def certified_degrees(net, positions, charges, k, n_max=16, tol=1e-6, margin=-0.25, n_samp=1024, headroom=2,
                      max_steps=3, precision="mixed", max_depth=None):
    """The network's per-level degrees (predict_p_level at `margin`), certified: one plan
    with tables `headroom` degrees higher evaluates at the predicted degrees and is checked
    against the FP64 direct sum on `n_samp` sampled targets; a miss raises every level by
    one (evaluate-time degrees, no new setup), at most `max_steps` times -- the plain
    fixed-degree climb as the network's judge.  Returns (degrees, error, climbs)."""
    from .api import evaluate, setup
    pl = predict_p_level(net, positions, charges, k, n_max=n_max, max_depth=max_depth, margin=margin)
    tab = [min(P_CAP, v + headroom) for v in pl]
    st = setup(positions, k, p=max(tab), N_max=n_max, p_level=tab, precision=precision, max_depth=max_depth)
    try:
        cur = list(pl)[:st.depth]
        st.set_degrees(cur)
        err = sampled_error(positions, charges, k, evaluate(st, charges), n_samp)
        climbs = 0
        while err > tol and climbs < max_steps:
            cur = [min(c + 1, t) for c, t in zip(cur, st.p_table)]
            st.set_degrees(cur)
            err = sampled_error(positions, charges, k, evaluate(st, charges), n_samp)
            climbs += 1
        return cur, err, climbs
    finally:
        st.close()


def evaluate_certified(tuned, plain, positions, charges, k, tol=1e-6, n_samp=256):
    """Evaluate with the tuned plan; if the sampled-target error misses `tol`, rerun
    with the plain fixed-degree plan.  Returns (phi, error, used_plain)."""
    from .api import evaluate
    phi = evaluate(tuned, charges)
    err = sampled_error(positions, charges, k, phi, n_samp)
    if err <= tol:
        return phi, err, False
    phi = evaluate(plain, charges)
    return phi, sampled_error(positions, charges, k, phi, n_samp), True
