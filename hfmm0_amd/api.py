"""Drop-in counterpart of the reference's ``HFMM.setup`` / ``HFMM.evaluate``.

``setup(positions, k, p=4, N_max=16, max_depth=None)`` and ``evaluate(state, charges)``
keep the reference signatures and ValueError messages (``HFMM.py:55-91``).  The
evaluate path runs entirely in libhfmm's HIP kernels; there is no CPU fallback.

Differences by design (the reference path is a kernel-independent FMM):
* ``p`` is the truncation degree of the solid-harmonic expansions ((p+1)^2
  coefficients per box), not the number of surface points per edge;
* ``evaluate`` also accepts complex charges (the reference accepts real ones only,
  ``HFMM.py:86-87``); real charges give the reference semantics, including an exactly
  real result when k == 0 (``HFMM.py:71-72, 90``).
"""
import ctypes

import numpy as np

from . import _lib as L

PRECISIONS = {"fp64": L.PREC_FP64, "mixed": L.PREC_MIXED}
# the automatic mixed-precision rule: M2L in FP32 where k D_l <= this (D_l the box
# diagonal of level l), FP64 above (hfmm.h: HFMM_MIXED_FP32_MAX_KD)
MIXED_FP32_MAX_KD = 32.0


class HFMMState:
    """Plan handle returned by :func:`setup`; owns all device buffers."""

    def __init__(self, handle, n, k, opts):
        self._h = handle
        self.n = n
        self.k = k
        self.opts = opts
        info = L.PlanInfo()
        L.check(L.lib().hfmm_plan_info_get(handle, ctypes.byref(info)))
        self.info = info

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("the HFMM state has been released")
        return self._h

    @property
    def depth(self):
        return self.info.depth

    @property
    def p_level(self):
        """Degree evaluated at levels 1..depth."""
        return [self.info.p_level[i] for i in range(1, self.info.depth + 1)]

    @property
    def m2l_fp32_levels(self):
        """Levels whose M2L runs in FP32 (mixed precision; empty at FP64)."""
        m = self.info.m2l_fp32_levels
        return [i for i in range(1, self.info.depth + 1) if (m >> i) & 1]

    @property
    def n_leaves_coord32(self):
        """Leaves whose near field takes FP32 coordinate differences (mixed precision)."""
        return int(self.info.n_leaves_coord32)

    @property
    def m2m_fp32_levels(self):
        """Levels l whose M2M (into l, from l+1) runs FP32 arithmetic (mixed precision)."""
        m = self.info.m2m_fp32_levels
        return [i for i in range(1, self.info.depth + 1) if (m >> i) & 1]

    @property
    def l2l_fp32_levels(self):
        """Levels l whose L2L (out of l, into l+1) runs FP32 arithmetic (mixed precision)."""
        m = self.info.l2l_fp32_levels
        return [i for i in range(1, self.info.depth + 1) if (m >> i) & 1]

    @property
    def p_table(self):
        """Degree the tables of levels 1..depth were built for."""
        return [self.info.p_table[i] for i in range(1, self.info.depth + 1)]

    def set_degrees(self, p_level):
        """Evaluate levels 1..depth at the given degrees (each <= its table degree; 0 or
        None keeps the table degree) without rebuilding the plan."""
        import numpy as np
        arr = np.zeros(L.MAX_LEVELS, dtype=np.int32)
        for i, v in enumerate(p_level):
            if i + 1 < L.MAX_LEVELS and v:
                arr[i + 1] = int(v)
        L.check(L.lib().hfmm_plan_set_degrees(self.handle, L.ptr(arr)))
        info = L.PlanInfo()
        L.check(L.lib().hfmm_plan_info_get(self.handle, ctypes.byref(info)))
        self.info = info

    def close(self):
        if self._h is not None:
            L.lib().hfmm_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tree(self):
        """Tree and lists of this plan as NumPy arrays (canonical box order)."""
        from .tree import _export
        lib = L.lib()
        return _export(self.handle, self.info, lib.hfmm_plan_export_tree, lib.hfmm_plan_export_list)


def setup(positions, k, p=4, N_max=16, max_depth=None, *, precision="fp64", p_level=None, use_graph=True,
          m2l_kernel="auto", tree_device=False, m2l_fp64_levels=None, p2p_tile=0, p2p_coord32_tol=0.0):
    """Build the octree, lists and operator tables; upload everything to the GPU.

    Validation order and messages follow ``HFMM.py:57-64`` and ``HFMM.py:14-16``.
    ``p_level`` optionally gives the truncation degree per tree level (index 0 is
    level 1); ``precision`` is ``"fp64"`` (plain FMM) or ``"mixed"`` (FP32 near field and
    FP32 M2L on matrix cores).  ``m2l_kernel="generic"`` forces the thread-per-coefficient
    M2L kernel instead of the matrix-core block kernel; an integer 8, 16, 32 or 64 asks the
    block kernel for that many pairs per operator application (target boxes times
    directions of equal theta and |d|; used where it fits).
    ``tree_device=True`` builds the octree and lists on the GPU (same arrays).
    ``m2l_fp64_levels`` (mixed precision only) lists the tree levels whose M2L stays in
    FP64; ``None`` applies the automatic rule (FP64 where k times the level's box diagonal
    exceeds ``MIXED_FP32_MAX_KD``).  ``p2p_tile`` (mixed precision) sets the targets per
    lane of the near-field kernel: 1, 2 or 4 (0: 2).  ``p2p_coord32_tol`` (mixed precision)
    is the estimated relative near-field error up to which a leaf takes FP32 coordinate
    differences (0: ``HFMM_P2P_COORD32_TOL``, 2^-15; negative: FP64 coordinates everywhere).
    """
    x, k = _check_problem(positions, k)
    o = make_opts(p, N_max, max_depth, precision=precision, p_level=p_level, use_graph=use_graph,
                  m2l_kernel=m2l_kernel, tree_device=tree_device, m2l_fp64_levels=m2l_fp64_levels,
                  p2p_tile=p2p_tile, p2p_coord32_tol=p2p_coord32_tol, _k=k)
    h = ctypes.c_void_p()
    L.check(L.lib().hfmm_plan_create(L.ptr(x), len(x), k, ctypes.byref(o), ctypes.byref(h)))
    return HFMMState(h, len(x), k, o)


def _check_problem(positions, k):
    x = np.asarray(positions, dtype=np.float64)
    if x.ndim != 2 or x.shape[1] != 3 or len(x) == 0 or not np.isfinite(x).all():
        raise ValueError("positions must be a nonempty finite (N,3) array")
    return np.ascontiguousarray(x), k


def make_opts(p=4, N_max=16, max_depth=None, *, precision="fp64", p_level=None, use_graph=True,
              m2l_kernel="auto", tree_device=False, m2l_fp64_levels=None, p2p_tile=0, p2p_coord32_tol=0.0,
              _k=0.0):
    """Validated plan options (the C ``hfmm_opts``), in the reference's validation order
    after the positions check (``HFMM.py:57-64``, ``HFMM.py:14-16``)."""
    if p < 2 or int(p) != p or N_max < 1 or int(N_max) != N_max:
        raise ValueError("p must be an integer >=2; N_max must be a positive integer")
    if max_depth is not None and (max_depth < 1 or int(max_depth) != max_depth):
        raise ValueError("max_depth must be a positive integer or None")
    if not np.isfinite(float(_k)):
        raise ValueError("k must be a finite real number")
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
    o = L.default_opts()
    o.p = int(p)
    o.n_max = int(N_max)
    o.max_depth = int(max_depth) if max_depth is not None else 0
    o.precision = PRECISIONS[precision]
    o.use_graph = 1 if use_graph else 0
    if m2l_kernel not in ("auto", "generic", 8, 16, 32, 64):
        raise ValueError("m2l_kernel must be 'auto', 'generic', 8, 16, 32 or 64")
    o.m2l_kernel = {"auto": 0, "generic": 1}.get(m2l_kernel, m2l_kernel)
    o.tree_device = 1 if tree_device else 0
    if p2p_tile not in (0, 1, 2, 4):
        raise ValueError("p2p_tile must be 0, 1, 2 or 4")
    o.p2p_tile = int(p2p_tile)
    if not np.isfinite(float(p2p_coord32_tol)):
        raise ValueError("p2p_coord32_tol must be a finite number")
    o.p2p_coord32_tol = float(p2p_coord32_tol)
    if m2l_fp64_levels is not None:
        lv = [int(v) for v in m2l_fp64_levels]
        if any(v < 1 or v >= L.MAX_LEVELS for v in lv):
            raise ValueError(f"m2l_fp64_levels entries must be tree levels in [1, {L.MAX_LEVELS - 1}]")
        o.m2l_fp64_levels = sum(1 << v for v in set(lv))
    if p_level is not None:
        for i, v in enumerate(p_level):
            if i + 1 >= L.MAX_LEVELS:
                break
            o.p_level[i + 1] = int(v)
    return o


def _charge_type(q):
    """(contiguous array, C type code) for real/complex charges of 32 or 64 bits."""
    if np.iscomplexobj(q):
        if q.dtype == np.complex64:
            return np.ascontiguousarray(q), L.C64
        return np.ascontiguousarray(q, dtype=np.complex128), L.C128
    if q.dtype == np.float32:
        return np.ascontiguousarray(q), L.F32
    return np.ascontiguousarray(q, dtype=np.float64), L.F64


def evaluate(state, charges, out_dtype=np.complex128):
    """phi at every point, input order (``HFMM.py:83-91``).  Real charges give the
    reference semantics; complex charges are accepted too.  float32 / complex64 charges
    are widened on the GPU.  ``out_dtype=np.complex64`` opts into a complex64 result
    (the FP64 result rounded on the GPU)."""
    q = np.asarray(charges)
    if q.shape != (state.n,):
        raise ValueError("charges must be a real array of shape (N,)")
    qc, qt = _charge_type(q)
    out_dtype = np.dtype(out_dtype)
    if out_dtype not in (np.complex128, np.complex64):
        raise ValueError("out_dtype must be complex128 or complex64")
    out = np.empty(state.n, dtype=out_dtype)
    ot = L.C128 if out_dtype == np.complex128 else L.C64
    L.check(L.lib().hfmm_evaluate_typed(state.handle, L.ptr(qc), qt, L.ptr(out), ot, L.HOST, None))
    return out


def evaluate_field(state, charges):
    """(phi, E) at every point, input order: phi (N,) complex128 and the field
    E = -grad phi (N, 3) complex128 -- the Helmholtz counterpart of the reference's
    ``kifmm.evaluate_setup(..., field=True)`` (``jaxfmm/kifmm/setup.py:229-261``)."""
    q = np.asarray(charges)
    if q.shape != (state.n,):
        raise ValueError("charges must be a real array of shape (N,)")
    qc, qt = _charge_type(q)
    phi = np.empty(state.n, dtype=np.complex128)
    field = np.empty((state.n, 3), dtype=np.complex128)
    L.check(L.lib().hfmm_evaluate_field(state.handle, L.ptr(qc), qt, L.ptr(phi), L.ptr(field), L.HOST, None))
    return phi, field


def torch_charge_code(dtype):
    """C type code (HFMM_F64 / C128 / F32 / C64) of a torch charge dtype; any other dtype
    raises ValueError."""
    import torch
    qt = {torch.float64: L.F64, torch.complex128: L.C128, torch.float32: L.F32,
          torch.complex64: L.C64}.get(dtype)
    if qt is None:
        raise ValueError("charges must be float32/64 or complex64/128")
    return qt


def device_charge_type(q_dev, out_dev, n):
    """C type code of a device-resident charge tensor, after checking both tensors: ``q_dev``
    (n,) float64 / complex128 / float32 / complex64 and ``out_dev`` (n,) complex128, both
    contiguous GPU tensors.  Every entry point that hands torch pointers to the library
    goes through this map, so the library never reads a buffer as a wider type than it is."""
    import torch
    if q_dev.shape != (n,) or out_dev.shape != (n,):
        raise ValueError("charges must be a real array of shape (N,)")
    if out_dev.dtype != torch.complex128:
        raise ValueError("the device output must be a complex128 tensor")
    qt = torch_charge_code(q_dev.dtype)
    if not (q_dev.is_cuda and out_dev.is_cuda and q_dev.is_contiguous() and out_dev.is_contiguous()):
        raise ValueError("device evaluate needs contiguous GPU tensors")
    return qt


def evaluate_device(state, q_dev, out_dev, stream=None):
    """Evaluate with device-resident torch tensors: ``q_dev`` (N,) float32/64 or
    complex64/128, ``out_dev`` (N,) complex128, both on the GPU."""
    import torch
    qt = device_charge_type(q_dev, out_dev, state.n)
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    L.check(L.lib().hfmm_evaluate(state.handle, ctypes.c_void_p(q_dev.data_ptr()), qt,
                                  ctypes.c_void_p(out_dev.data_ptr()), L.DEVICE, ctypes.c_void_p(stream)))
    return out_dev


def benchmark(state, q_dev, out_dev, warmup=3, reps=20, phases=False):
    """Median milliseconds of one device-resident evaluate (hipEvent timing of graph
    replays) and optionally per-phase milliseconds of one instrumented run."""
    qt = device_charge_type(q_dev, out_dev, state.n)
    med = ctypes.c_double()
    ph = np.zeros(7)
    L.check(L.lib().hfmm_benchmark(state.handle, ctypes.c_void_p(q_dev.data_ptr()), qt,
                                   ctypes.c_void_p(out_dev.data_ptr()), int(warmup), int(reps),
                                   ctypes.byref(med), L.ptr(ph) if phases else None))
    if phases:
        names = ["gather", "p2m", "m2m", "m2l", "l2l", "l2p", "p2p"]
        return med.value, dict(zip(names, ph.tolist()))
    return med.value


def benchmark_m2l_levels(state, q_dev, out_dev, reps=5):
    """Average milliseconds of each level's M2L launch ({level: ms}), hipEvents around
    every launch on the plan's stream over `reps` instrumented evaluates."""
    qt = device_charge_type(q_dev, out_dev, state.n)
    ms = np.zeros(L.MAX_LEVELS)
    L.check(L.lib().hfmm_benchmark_m2l_levels(state.handle, ctypes.c_void_p(q_dev.data_ptr()), qt,
                                              ctypes.c_void_p(out_dev.data_ptr()), int(reps), L.ptr(ms)))
    return {lv: float(ms[lv]) for lv in range(L.MAX_LEVELS) if ms[lv] > 0}


def direct(positions, charges, k, targets):
    """FP64 direct sum on the GPU at ``positions[targets]`` (sampled-target oracle)."""
    x = np.ascontiguousarray(positions, dtype=np.float64)
    q = np.ascontiguousarray(np.asarray(charges).astype(np.complex128))
    t = np.ascontiguousarray(targets, dtype=np.int64)
    out = np.empty(len(t), dtype=np.complex128)
    L.check(L.lib().hfmm_direct(L.ptr(x), len(x), L.ptr(q), float(k), L.ptr(t), len(t), L.ptr(out)))
    return out
