"""Oracle (test infrastructure only): the leveled-KIFMM restatement of the reference's
evaluate (oracle/kifmm.py) with its evaluate phases in C/OpenMP (oracle/kifmm_eval.c),
so that the reference's algorithm can be timed on host cores with the setup excluded.

``prepare`` does the setup of oracle/kifmm.py (tree, lists, surfaces, Tikhonov-factored
per-level operators for the cos and sin kernels, HFMM.py:55-80); ``evaluate`` runs the two
components' phases in C and returns real + 1j imag in input order (HFMM.py:83-91).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import it.  Pinned by tests/test_oracle_kifmm.py: equal to oracle/kifmm.py's NumPy
evaluate to rounding on the same inputs.
"""
import ctypes
import os
import subprocess

import numpy as np

from oracle import kifmm as K
from oracle import tree as T

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_P = ctypes.c_void_p


class _Args(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("nb", ctypes.c_int), ("ns", ctypes.c_int), ("depth", ctypes.c_int),
                ("trig", ctypes.c_int), ("k", ctypes.c_double),
                ("x", _P), ("q", _P), ("level", _P), ("pstart", _P), ("count", _P), ("octant", _P),
                ("is_leaf", _P), ("centre", _P), ("h", _P), ("eq", _P), ("ck", _P),
                ("fup_u", _P), ("fup_d", _P), ("fup_vt", _P), ("fdn_u", _P), ("fdn_d", _P), ("fdn_vt", _P),
                ("m2m", _P), ("l2l", _P), ("m2l", _P), ("child_ptr", _P), ("child", _P),
                ("n_vclass", ctypes.c_int), ("vclass_ptr", _P), ("vclass_key", _P), ("v_tgt", _P), ("v_src", _P),
                ("u_ptr", _P), ("u_src", _P), ("w_ptr", _P), ("w_src", _P), ("x_ptr", _P), ("x_src", _P)]


def build(force=False):
    so = os.path.join(_HERE, "liboracle_kifmm.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", _HERE, "liboracle_kifmm.so"], check=True)
    return so


def _lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle_kifmm.so")
        if not os.path.exists(so):
            build()
        _LIB = ctypes.CDLL(so)
        _LIB.oracle_kifmm_eval.restype = ctypes.c_int
        _LIB.oracle_kifmm_eval.argtypes = [ctypes.POINTER(_Args), _P]
    return _LIB


def _csr(tgt, src, nb):
    tgt = np.asarray(tgt, np.int64)
    order = np.argsort(tgt, kind="stable")
    ptr = np.zeros(nb + 1, np.int32)
    np.add.at(ptr, tgt + 1, 1)
    return np.cumsum(ptr).astype(np.int32), np.asarray(src, np.int32)[order]


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Prepared:
    """Setup of oracle/kifmm.py for one point set, k and reference order p."""

    def __init__(self, positions, k, p=4, n_max=16, max_depth=None):
        x = np.asarray(positions, dtype=np.float64)
        self.k = float(k)
        tr = T.build_tree(x, n_max, max_depth)
        lists = T.build_lists(tr)
        centres = T.box_centres(tr)
        h = tr["extent"] * 0.5 ** (tr["level"].astype(np.float64) + 1.0)
        nb, md = len(tr["level"]), tr["depth"]
        self.perm = tr["perm"].astype(np.int64)
        self.n = len(x)
        self.arrays = dict(x=_c(x[self.perm], np.float64), level=_c(tr["level"], np.int32),
                           pstart=_c(tr["pstart"], np.int32), count=_c(tr["count"], np.int32),
                           octant=_c(np.asarray(tr["code"], np.uint64) & np.uint64(7), np.int32),
                           is_leaf=_c(tr["is_leaf"], np.uint8), centre=_c(centres, np.float64), h=_c(h, np.float64),
                           eq=_c(K.surface_pts(p, K.R_EQUIV), np.float64), ck=_c(K.surface_pts(p, K.R_CHECK), np.float64))
        ns = len(self.arrays["eq"])
        parent = np.asarray(tr["parent"])
        kids = np.nonzero(parent >= 0)[0]
        self.arrays["child_ptr"], self.arrays["child"] = _csr(parent[kids], kids, nb)
        V = np.asarray(lists["V"]).reshape(-1, 2)
        gt = np.stack(T.morton_grid(tr["code"][V[:, 0]]), 1).astype(np.int64)
        gs = np.stack(T.morton_grid(tr["code"][V[:, 1]]), 1).astype(np.int64)
        d = gt - gs
        key = tr["level"][V[:, 0]].astype(np.int64) * 343 + (d[:, 0] + 3) * 49 + (d[:, 1] + 3) * 7 + (d[:, 2] + 3)
        order = np.argsort(key, kind="stable")
        keys, starts = np.unique(key[order], return_index=True)
        self.arrays["vclass_ptr"] = _c(np.append(starts, len(order)), np.int32)
        self.arrays["vclass_key"] = _c(keys, np.int32)
        self.arrays["v_tgt"] = _c(V[order, 0], np.int32)
        self.arrays["v_src"] = _c(V[order, 1], np.int32)
        self.n_vclass = len(keys)
        for name in ("U", "W", "X"):
            L = np.asarray(lists[name]).reshape(-1, 2)
            self.arrays[name.lower() + "_ptr"], self.arrays[name.lower() + "_src"] = _csr(L[:, 0], L[:, 1], nb)
        self.nb, self.ns, self.depth = nb, ns, md
        # per-level operators of the two kernels (eps_rel 1e-7 real, 1e-12 imaginary, HFMM.py:75-77)
        self.ops = []
        comps = [(0, np.cos, 1e-7)] + ([] if self.k == 0.0 else [(1, np.sin, 1e-12)])
        for trig, fn, eps in comps:
            lv = [K.level_operators(p, tr["extent"] * 0.5 ** (l + 1), self.k, fn, eps) for l in range(md + 1)]
            o = {}
            for tag, f in (("up", "F_up"), ("dn", "F_dn")):
                o["f%s_u" % tag] = _c(np.stack([op[f][0] for op in lv]), np.float64)
                o["f%s_d" % tag] = _c(np.stack([op[f][1] for op in lv]), np.float64)
                o["f%s_vt" % tag] = _c(np.stack([op[f][2] for op in lv]), np.float64)
            o["m2m"] = _c(np.stack([op["M2M_K"] for op in lv]), np.float64)
            o["l2l"] = _c(np.stack([op["L2L_K"] for op in lv]), np.float64)
            m2l = np.zeros((md + 1, 343, ns, ns))
            for l, op in enumerate(lv):
                for (dx, dy, dz), mat in op["M2L_K"].items():
                    m2l[l, (dx + 3) * 49 + (dy + 3) * 7 + (dz + 3)] = mat
            o["m2l"] = m2l
            self.ops.append((trig, o))

    def _component(self, trig, o, q_sorted):
        a = _Args()
        a.n, a.nb, a.ns, a.depth, a.trig, a.k = self.n, self.nb, self.ns, self.depth, trig, self.k
        a.q = q_sorted.ctypes.data
        for name, arr in list(self.arrays.items()) + list(o.items()):
            setattr(a, name, arr.ctypes.data)
        a.n_vclass = self.n_vclass
        pot = np.empty(self.n)
        if _lib().oracle_kifmm_eval(ctypes.byref(a), pot.ctypes.data) != 0:
            raise MemoryError("oracle_kifmm_eval: allocation failed")
        return pot

    def evaluate(self, charges):
        """real + 1j imag for real charges (reference semantics, HFMM.py:83-91); complex
        charges by linearity."""
        q = np.asarray(charges)
        if np.iscomplexobj(q):
            return self.evaluate(q.real) + 1j * self.evaluate(q.imag)
        qs = _c(q.astype(np.float64)[self.perm], np.float64)
        out = np.empty(self.n, dtype=np.complex128)
        parts = [self._component(trig, o, qs) for trig, o in self.ops]
        res = parts[0] + (1j * parts[1] if len(parts) > 1 else 0.0)
        out[self.perm] = res
        return out
