"""ctypes binding of libhfmm.so (the C ABI in include/hfmm.h).

The library is loaded from this package directory only; if it is missing the import
fails loudly -- there is no fallback implementation.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhfmm.so")

HFMM_OK = 0
HFMM_ERR_ARG = -2
HFMM_ERR_DEVICE = -3
HFMM_ERR_MEMORY = -4
PREC_FP64 = 0
PREC_MIXED = 1
HOST = 0
DEVICE = 1
F64 = 0
C128 = 1
F32 = 2
C64 = 3
MAX_LEVELS = 24


class Opts(ctypes.Structure):
    _fields_ = [("p", ctypes.c_int), ("p_level", ctypes.c_int * MAX_LEVELS), ("n_max", ctypes.c_int),
                ("max_depth", ctypes.c_int), ("precision", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("m2l_kernel", ctypes.c_int), ("tree_device", ctypes.c_int),
                ("m2l_fp64_levels", ctypes.c_int), ("p2p_tile", ctypes.c_int),
                ("p2p_coord32_tol", ctypes.c_double)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("n_points", ctypes.c_int64), ("depth", ctypes.c_int), ("n_boxes", ctypes.c_int),
                ("n_leaves", ctypes.c_int), ("maxleaf", ctypes.c_int), ("n_v", ctypes.c_int64),
                ("n_u", ctypes.c_int64), ("n_w", ctypes.c_int64), ("n_x", ctypes.c_int64),
                ("n_p2p_pairs", ctypes.c_int64), ("p_level", ctypes.c_int * MAX_LEVELS),
                ("extent", ctypes.c_double), ("pmin", ctypes.c_double * 3),
                ("p_table", ctypes.c_int * MAX_LEVELS), ("m2l_fp32_levels", ctypes.c_int),
                ("n_leaves_phase64", ctypes.c_int), ("m2m_fp32_levels", ctypes.c_int),
                ("l2l_fp32_levels", ctypes.c_int), ("n_leaves_coord32", ctypes.c_int)]


class HFMMError(RuntimeError):
    pass


_lib = None
vp = ctypes.c_void_p


def _declare(lib):
    I, D, L = ctypes.c_int, ctypes.c_double, ctypes.c_int64
    sig = {
        "hfmm_opts_default": (None, [ctypes.POINTER(Opts)]),
        "hfmm_plan_create": (I, [vp, L, D, ctypes.POINTER(Opts), ctypes.POINTER(vp)]),
        "hfmm_plan_info_get": (I, [vp, ctypes.POINTER(PlanInfo)]),
        "hfmm_plan_export_tree": (I, [vp, vp, vp, vp, vp, vp, vp, vp]),
        "hfmm_plan_export_list": (I, [vp, ctypes.c_char, vp]),
        "hfmm_evaluate": (I, [vp, vp, I, vp, I, vp]),
        "hfmm_evaluate_typed": (I, [vp, vp, I, vp, I, I, vp]),
        "hfmm_evaluate_field": (I, [vp, vp, I, vp, vp, I, vp]),
        "hfmm_plan_set_degrees": (I, [vp, vp]),
        "hfmm_benchmark": (I, [vp, vp, I, vp, I, I, ctypes.POINTER(D), vp]),
        "hfmm_benchmark_m2l_levels": (I, [vp, vp, I, vp, I, vp]),
        "hfmm_plan_destroy": (I, [vp]),
        "hfmm_batch_create": (I, [I, vp, vp, vp, vp, ctypes.POINTER(vp)]),
        "hfmm_batch_plan": (I, [vp, I, ctypes.POINTER(vp)]),
        "hfmm_batch_evaluate": (I, [vp, vp, I, vp, I, vp]),
        "hfmm_batch_benchmark": (I, [vp, vp, I, vp, I, I, ctypes.POINTER(D)]),
        "hfmm_batch_benchmark_phases": (I, [vp, vp, I, vp, I, vp]),
        "hfmm_batch_destroy": (I, [vp]),
        "hfmm_direct": (I, [vp, L, vp, D, vp, L, vp]),
        "hfmm_direct_noalloc": (I, [vp, L, vp, D, vp, L, vp, vp]),
        "hfmm_tree_create": (I, [vp, L, I, I, ctypes.POINTER(vp)]),
        "hfmm_tree_create_device": (I, [vp, L, I, I, ctypes.POINTER(vp)]),
        "hfmm_tree_info_get": (I, [vp, ctypes.POINTER(PlanInfo)]),
        "hfmm_tree_export": (I, [vp, vp, vp, vp, vp, vp, vp, vp]),
        "hfmm_tree_export_list": (I, [vp, ctypes.c_char, vp]),
        "hfmm_tree_destroy": (I, [vp]),
        "hfmm_host_jhat": (I, [D, I, vp]),
        "hfmm_host_hhat": (I, [D, I, vp, vp]),
        "hfmm_host_rotation": (I, [D, I, vp]),
        "hfmm_host_coax": (I, [I, D, D, D, D, I, I, vp]),
        "hfmm_mlp_create": (I, [vp, ctypes.POINTER(vp)]),
        "hfmm_mlp_params": (I, [vp, vp]),
        "hfmm_mlp_forward": (I, [vp, vp, L, vp]),
        "hfmm_mlp_gradient": (I, [vp, vp, vp, L, vp, vp]),
        "hfmm_mlp_train": (I, [vp, vp, vp, L, I, I, ctypes.c_float, I, vp]),
        "hfmm_mlp_destroy": (I, [vp]),
        "hfmm_last_error": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return sig


EXPORTED = [
    "hfmm_opts_default", "hfmm_plan_create", "hfmm_plan_info_get", "hfmm_plan_export_tree", "hfmm_plan_set_degrees",
    "hfmm_plan_export_list", "hfmm_evaluate", "hfmm_evaluate_typed", "hfmm_evaluate_field", "hfmm_benchmark", "hfmm_benchmark_m2l_levels",
    "hfmm_plan_destroy", "hfmm_direct", "hfmm_direct_noalloc",
    "hfmm_tree_create", "hfmm_tree_create_device", "hfmm_tree_info_get", "hfmm_tree_export", "hfmm_tree_export_list",
    "hfmm_tree_destroy", "hfmm_host_jhat", "hfmm_host_hhat", "hfmm_host_rotation", "hfmm_host_coax",
    "hfmm_mlp_create", "hfmm_mlp_params", "hfmm_mlp_forward", "hfmm_mlp_gradient", "hfmm_mlp_train",
    "hfmm_mlp_destroy", "hfmm_batch_create", "hfmm_batch_plan", "hfmm_batch_evaluate", "hfmm_batch_benchmark", "hfmm_batch_benchmark_phases",
    "hfmm_batch_destroy", "hfmm_last_error",
]


def lib():
    """The loaded library; raises if libhfmm.so has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HFMMError(f"{LIB_PATH} is missing: build it with `python -m hfmm0_amd.build` "
                            "(there is no non-HIP implementation)")
        # torch (ROCm build) ships its own libamdhip64.so with the soname libamdhip64.so.7.
        # Loading torch first makes libhfmm bind to that already-loaded runtime, so a
        # process that also uses torch holds a single HIP runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        l = ctypes.CDLL(LIB_PATH)
        _declare(l)
        _lib = l
    return _lib


def check(status):
    """Map a status code to an exception (ValueError for argument errors)."""
    if status == HFMM_OK:
        return
    msg = lib().hfmm_last_error().decode()
    if status == HFMM_ERR_ARG:
        raise ValueError(msg)
    if status == HFMM_ERR_MEMORY:
        raise MemoryError(msg)
    raise HFMMError(f"libhfmm status {status}: {msg}")


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if isinstance(a, np.ndarray) else ctypes.c_void_p(a)


def default_opts():
    o = Opts()
    lib().hfmm_opts_default(ctypes.byref(o))
    return o
