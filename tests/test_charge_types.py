"""Host checks of the torch dtype -> C charge type map used by every device-pointer entry
point (evaluate_device, benchmark, benchmark_m2l_levels, Batch): each dtype maps to the
type of its own width, and anything else is rejected before a pointer reaches the library."""
import pytest
import torch

from hfmm0_amd import _lib as L
from hfmm0_amd.api import device_charge_type, torch_charge_code


@pytest.mark.parametrize("dtype,code", [(torch.float64, L.F64), (torch.complex128, L.C128),
                                        (torch.float32, L.F32), (torch.complex64, L.C64)])
def test_charge_code_matches_width(dtype, code):
    assert torch_charge_code(dtype) == code


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.int64, torch.int32, torch.uint8])
def test_other_dtypes_rejected(dtype):
    with pytest.raises(ValueError):
        torch_charge_code(dtype)


def test_output_must_be_complex128():
    q = torch.zeros(5, dtype=torch.float32)
    for od in (torch.complex64, torch.float64):
        with pytest.raises(ValueError, match="complex128"):
            device_charge_type(q, torch.zeros(5, dtype=od), 5)


def test_shape_and_device_checked():
    with pytest.raises(ValueError, match="shape"):
        device_charge_type(torch.zeros(4), torch.zeros(5, dtype=torch.complex128), 5)
    with pytest.raises(ValueError, match="GPU"):
        device_charge_type(torch.zeros(5), torch.zeros(5, dtype=torch.complex128), 5)
