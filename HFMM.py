"""Drop-in module name of the reference API: ``from HFMM import setup, evaluate``.

The implementation is ``hfmm0_amd.api`` (HIP kernels through libhfmm's C ABI).
"""
from hfmm0_amd.api import evaluate, setup  # noqa: F401

__all__ = ["setup", "evaluate"]
