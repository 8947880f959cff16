"""MI355X-native Helmholtz FMM (HIP kernels for gfx950 behind a C ABI).

``from hfmm0_amd import setup, evaluate`` mirrors the reference's ``HFMM`` module.
"""
from .api import (HFMMState, benchmark, benchmark_m2l_levels, direct, evaluate, evaluate_device,  # noqa: F401
                  evaluate_field, setup)
