"""Host octree and interaction lists built by libhfmm (no device use).

Mirrors the reference's ``build_fmm_tree_auto`` + ``assign_particles``
(``jaxfmm/trees/cubic.py:530-695``) and list builders (``jaxfmm/trees/lists.py:818-931``)
with boxes in canonical order (level ascending, Morton code ascending).
"""
import ctypes

import numpy as np

from . import _lib as L


def _export(handle, info, export_tree, export_list):
    nb, n = info.n_boxes, info.n_points
    out = dict(level=np.empty(nb, np.int32), code=np.empty(nb, np.uint64), count=np.empty(nb, np.int32),
               pstart=np.empty(nb, np.int32), parent=np.empty(nb, np.int32), is_leaf=np.empty(nb, np.uint8),
               perm=np.empty(n, np.int64))
    L.check(export_tree(handle, *(L.ptr(out[k]) for k in ("level", "code", "count", "pstart", "parent",
                                                          "is_leaf", "perm"))))
    out["is_leaf"] = out["is_leaf"].astype(bool)
    for which, cnt in (("V", info.n_v), ("U", info.n_u), ("W", info.n_w), ("X", info.n_x)):
        a = np.empty((cnt, 2), np.int32)
        if cnt:
            L.check(export_list(handle, which.encode(), L.ptr(a)))
        out[which] = a
    out["depth"] = info.depth
    out["extent"] = info.extent
    out["pmin"] = np.array(list(info.pmin))
    out["maxleaf"] = info.maxleaf
    return out


def build_tree(positions, n_max=16, max_depth=None, device=False):
    """Tree and lists as NumPy arrays (dict) from the C++ host builder, or (device=True)
    from the GPU builder, whose arrays are identical."""
    x = np.ascontiguousarray(positions, dtype=np.float64)
    if x.ndim != 2 or x.shape[1] != 3:
        raise ValueError("positions must be a nonempty finite (N,3) array")
    lib = L.lib()
    h = ctypes.c_void_p()
    create = lib.hfmm_tree_create_device if device else lib.hfmm_tree_create
    L.check(create(L.ptr(x), len(x), int(n_max), int(max_depth or 0), ctypes.byref(h)))
    try:
        info = L.PlanInfo()
        L.check(lib.hfmm_tree_info_get(h, ctypes.byref(info)))
        return _export(h, info, lib.hfmm_tree_export, lib.hfmm_tree_export_list)
    finally:
        lib.hfmm_tree_destroy(h)
